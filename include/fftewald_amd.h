/*
 * fftewald_amd.h -- C ABI of the MI355X (gfx950) periodic Coulomb engine.
 *
 * Drop-in for the Ewald library of the DFTB+ fast-Ewald fork, whose public surface is
 *   ewald_potential_force, ewald_parameters, ewald_grid_memory, TEwaldParameters3d
 *   (src/dftbp/extlibs/fftewald/fft/ewald_fft_3d.f90:19-20)
 * and for the per-SCC-iteration calls TCoulomb makes into it
 *   (src/dftbp/dftb/coulomb.F90:336-385 fftEwaldShift / fftEwaldAddGradients,
 *    786-880 updateShifts), composed with the short-range gamma of TShortGamma
 *   (src/dftbp/dftb/shortgamma.F90:860-934) as TScc%updateShifts does (scc.F90:453-496).
 *
 * Plain C: pointers and sizes only, no HIP or torch types.  Every call returns an int status
 * (EW_OK = 0); none aborts.  ew_last_error() gives the message of the last failure.
 *
 * Units are atomic (Bohr, e, Hartree), as in the reference (fft3d_readme.md:17).
 * Array layouts are DFTB+'s native ones, so the Fortran caller passes its arrays unchanged:
 *   coords   [n][3]  row i = Cartesian position of atom i      (DFTB+ coords(3,N))
 *   latvecs  [3][3]  row i = lattice vector a_i                 (DFTB+ latVecs(3,3) columns)
 *   dq, shift, pot   [n]
 *   force, grad      [n][3]
 * The reference's own ewald_potential_force takes positions(N,3) and latVecs with rows a_i in
 * Fortran order; fortran/fftewald_amd.f90 provides that exact signature on top of this ABI.
 *
 * Threading: one handle per host thread; each handle owns a HIP stream on its device.
 */
#ifndef FFTEWALD_AMD_H
#define FFTEWALD_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define EW_OK 0
#define EW_ERR_ARGUMENT 1     /* null pointer, n <= 0, bad option value                      */
#define EW_ERR_NOT_NEUTRAL 2  /* |sum q| > 1e-8 sum |q|  (fe/common/ewald_validation.f90:41-46) */
#define EW_ERR_CELL 3         /* cell volume <= 1e-18    (fe/common/ewald_geometry.f90:78)      */
#define EW_ERR_TOLERANCE 4    /* tol outside (0,1)       (fe/fft/ewald_fft_3d.f90:97-98)        */
#define EW_ERR_COINCIDENT 5   /* two charges closer than 1e-12 Bohr
                                 (fe/fft/ewald_fft_3d_real.f90:159-160)                     */
#define EW_ERR_NO_GEOMETRY 6  /* ew_shift before ew_set_geometry                                 */
#define EW_ERR_DEVICE 7       /* HIP runtime failure                                              */
#define EW_ERR_MEMORY 8       /* device memory allocation failed / over the admission budget     */
#define EW_ERR_NOT_FINITE 9   /* NaN or Inf in coordinates or charges                             */
#define EW_ERR_PENDING 10     /* ew_server_wait: the request is not finished within the
                                 timeout                                                    */

/* precision modes */
#define EW_PRECISION_FP64 0   /* every step in FP64                                               */
#define EW_PRECISION_MIXED 1  /* FP32 spread/FFT/gather/pair coefficients, FP64 accumulation,
                                 self term and energy                                             */

typedef struct ew_handle ew_handle;

/* Launch shapes of one structure's evaluation: how the spreading is blocked and how the 3D FFT
 * is cut into passes.  Chosen by fixed rules, by the autotuner (ew_options.tune) or given.  No
 * counterpart in the reference, whose analogue is FFTW's per-call planning
 * (fe/nfft/fft_backend.f90:52-56). */
typedef struct ew_tuning {
  int spread_direct;    /* 2: owned blocks -- each spreading block takes every atom whose stencil
                           reaches it and stores its own mesh points once (no copies);
                           1: each block adds its LDS copy (block + halo) into the cleared mesh
                           with device atomics; 0: copies to scratch, then a fixed-order
                           assembly                                                          */
  int spread_block[3];  /* SPME sort tiles (16x8x8 mesh points) per spreading block per axis;
                           0 0 0: no privatised blocks                                         */
  int fft_planes;       /* 1: fused x-y plane transforms (a whole plane per workgroup)          */
  int fft_lines[3];     /* lines per workgroup of the x (pairs of real rows), y and z passes   */
  int fft_radix[3];     /* radix order per axis: 0 = 8s, 4, 2, then 3s, 5s; 1 = 4s, 2, 3s, 5s;
                           2 = the leftover power of two first, then 8s, 3s, 5s;
                           3 = 3s and 5s first, then 8s, 4, 2; 4 = 16s, then 8, 4 or 2, then
                           3s, 5s                                                           */
  int pair_cosched;     /* 1: the real-space pair sum runs on extra workgroups of the spreading
                           and FFT launches, on the evaluation stream; 0: on an auxiliary
                           stream beside them (fork and join by events)                       */
  int pair_spread_share; /* pair_cosched: percent of the pair rows given to the spreading
                           launch; the FFT launches share the rest equally                    */
  float eval_us;        /* autotuned: the measured evaluation time of this choice; else 0       */
  int fft_wave;         /* register transforms in place of the Stockham passes above (a DFT of
                           length K/lanes per lane, one LDS exchange per line, compile-time index
                           arithmetic): 2 = x-y plane kernels (square planes of side 32, 64, 96
                           or 128, one workgroup per z plane) and a z line pass; 1 = x, y and z
                           line passes (one wave per workgroup; axis lengths 32, 48, 64, 96, 128,
                           256); a shape without a compiled kernel falls back to 1, then 0;
                           0 = the Stockham passes                                          */
  int fft_lanes[3];     /* fft_wave: lanes per line of the x, y and z transforms (4, 8 or 16;
                           the plane kernels use fft_lanes[0] for x and y)                   */
  int pair_blocks;      /* workgroups of the real-space pair sum when it runs beside the chain
                           on the auxiliary stream (fewer leave the chain more CUs and HBM
                           bandwidth); 0: 512                                                 */
  int pair_fork;        /* 0: the pair sum starts with the evaluation, beside the spreading;
                           1: after the spreading, beside the FFTs and the mesh gather        */
} ew_tuning;

/* Options; zero-initialise and set what you need.  Fields <= 0 mean "choose automatically",
 * mirroring the optional *_in overrides of ewald_potential_force
 * (fe/fft/ewald_fft_3d.f90:36-76). */
typedef struct ew_options {
  double tol;      /* requested accuracy in (0,1); 0 -> 1e-10 (fe/fft/ewald_fft_3d.f90:23)       */
  double alpha;    /* splitting parameter override (alpha_in), Bohr^-1                           */
  double r_cut;    /* real-space cutoff override (r_cut_in), Bohr                                */
  int grid[3];     /* SPME grid K1,K2,K3 override (the mesh analogue of nModes_in)               */
  int order;       /* B-spline order override (the stencil analogue of windowCutoff_in), 2..16  */
  int precision;   /* EW_PRECISION_FP64 (default) or EW_PRECISION_MIXED                          */
  int device;      /* HIP device ordinal                                                          */
  int alpha_rule;  /* automatic alpha: 0 = the fast path's real-space budget rule
                      (fe/fft/ewald_fft_3d_parameters.f90:114-121); 1 = DFTB+'s
                      getOptimalAlphaEwald(tol) (src/dftbp/dftb/coulomb.F90:2097-2179) with
                      r_cut = sqrt(-ln tol)/alpha (fe/common/ewald_truncation.f90:36-45)     */
  int profile;     /* 1: record HIP events around every phase of each evaluation, every phase alone
                      on the handle's stream (ew_phase_times); 2: record HIP events in the
                      evaluation's own launch shape -- pair sum beside the chain, direct
                      launches -- without changing it (ew_trace_times)                       */
  int eager;       /* 1: launch the evaluation's kernels one by one; 0 (default): replay a HIP
                      graph captured once per (dq, shift, energy) buffer triple, except for
                      large structures (>= 32768 atoms with privatised spreading), whose
                      direct launches measured faster than replay                             */
  int charged;     /* 1: accept non-neutral charges and add the neutralising background
                      -pi/(V alpha^2) * sum(q) to every shift, as DFTB+'s stock Ewald does
                      (src/dftbp/dftb/coulomb.F90:1292-1296); 0 (default): reject them, as the
                      fast path does (coulomb.F90:255-265, fe/common/ewald_validation.f90:26-48) */
  int fft_mode;    /* 0 (default): the register x-y plane kernels where the plane is square with
                      a compiled side, else register line passes where every axis has a
                      compiled length, else the Stockham passes (fused planes in mixed
                      precision where a plane fits one workgroup's LDS); 1: always separate x
                      and y passes; 2: fused planes wherever a plane fits                  */
  int pair_storage; /* 0 (default): the row list (every linked pair in both rows, the fastest
                       pair sum); 1: each linked pair stored once (symmetric, tile-blocked):
                       half the pair memory, about twice the pair-sum time; its transposed
                       half adds through atomics, so results repeat to round-off only         */
  int group_grid;  /* grouped launches (pools, several charge vectors): 0 (default): one
                      workgroup per work item (structure, block); 1: persistent grids of at
                      most 4 workgroups per CU that walk the table of work items            */
  double k_cut;    /* Fourier cutoff override (k_cut_in), Bohr^-1: replaces the cutoff of the
                      Fourier budget (fe/fft/ewald_fft_3d_parameters.f90:127-130) and the mesh
                      follows from it as the modes follow from k_cut in modes_along_axis
                      (:159-167): K_i = smallest 2^a 3^b 5^c >= 2 sigma k_cut |a_i| / 2 pi;
                      an explicit grid still wins, as nModes_in does                          */
  int tune;        /* launch shapes (ew_tuning): 0 (default): fixed rules; 1: autotune -- at
                      ew_set_geometry, time candidate shapes on the device with the structure's
                      own data and keep the fastest, once per (mesh, order, precision, size
                      class, pair storage, device), persisted in a cache file
                      ($EW_TUNE_CACHE, else ewald_tune.cache beside the library); 2: cached
                      choices only (fixed rules on a miss); 3: `tuning` as given.  Results
                      repeat bit for bit with owned spreading blocks (spread_direct 2, which
                      the rules choose wherever a block fits the LDS); the block-copy and
                      direct-atomic write-outs add floating-point values through atomics
                      and repeat to round-off only                                        */
  ew_tuning tuning; /* tune = 3: the shapes to use                                              */
} ew_options;

/* Parameters an evaluation uses; mirrors TEwaldParameters3d
 * (fe/fft/ewald_fft_3d_parameters.f90:25-45) with the SPME mesh in place of the NFFT modes. */
typedef struct ew_parameters {
  double alpha;      /* splitting parameter                                                      */
  double r_cut;      /* Ewald real-space cutoff                                                  */
  double k_cut;      /* Fourier cutoff the mesh resolves, as a physical wavenumber              */
  double r_list;     /* radius of the resident pair list: max(r_cut, short-gamma cutoffs)        */
  int grid[3];       /* SPME mesh                                                                */
  int order;         /* B-spline order                                                           */
  int precision;
  int pair_mode;     /* 0 = linked tiles (minimum image), each pair in both rows;
                        1 = dense image sum (small cells); 2 = linked tiles, each pair stored
                        once (symmetric, tile-blocked)                                         */
  int index_bytes;   /* bytes per stored partner index: 2 (offsets from the row) or 4          */
  int fft_planes;    /* 1: fused x-y plane transforms (a whole plane in one workgroup's LDS,
                        register or Stockham transforms); 0: separate x and y passes           */
  int gather_split;  /* 1: the pair sum runs beside the chain and the gather splits around the
                        join (mesh potentials before it, the shift and energy after it)       */
  long long n_pairs; /* stored real-space pair coefficients                                      */
  long long device_bytes; /* device memory held by the handle for this geometry                 */
} ew_parameters;

/* ---- resident engine (TCoulomb + TShortGamma state, kept on the device) ---- */

/* Create an engine.  opts may be NULL (all defaults).  No geometry yet. */
int ew_create(const ew_options *opts, ew_handle **out);
int ew_destroy(ew_handle *h);

/* Message of the last failure on h (or of the last failed ew_create / one-shot call when
 * h is NULL).  Never NULL. */
const char *ew_last_error(const ew_handle *h);

/* Geometry step: TCoulomb%updateLatVecs + updateCoords and TShortGamma%updateCoords
 * (coulomb.F90:509-620, shortgamma.F90:177-195).  Builds everything that depends only on
 * the geometry and keeps it on the device: atom ordering, B-spline weights, the pair list
 * with its erfc(alpha r)/r - gamma_short coefficients, FFT tables.
 *   species   [n] in 0..n_species-1, hubbard_u [n_species] (Hartree): when both are given the
 *             short-range gamma of shortgamma.F90 is folded into the pair coefficients and the
 *             returned shift is the SCC shift (Coulomb + short gamma); NULL -> Coulomb only.
 *   charges   optional [n]: only sum q^2 is read, for the automatic alpha (as
 *             ewald_parameters does); NULL -> a unit-charge estimate.
 * A later call replaces the geometry (n, cell and species may all change, as along a
 * trajectory): the previous device buffers are reused where the sizes fit, and every captured
 * evaluation is dropped, so the caller's buffers may be passed again to the entry points below. */
int ew_set_geometry(ew_handle *h, int n, const double *coords, const double *latvecs,
                    const int *species, const double *hubbard_u, int n_species,
                    const double *charges);

/* Parameters of the current geometry (ewald_parameters, fe/fft/ewald_fft_3d.f90:142-191). */
int ew_get_parameters(const ew_handle *h, ew_parameters *out);

/* The launch shapes the handle's evaluations use for its current geometry. */
int ew_get_tuning(const ew_handle *h, ew_tuning *out);

/* One SCC iteration, host arrays: shift = phi(dq) and energy = 1/2 sum shift*dq
 * (coulomb.F90:786-880, energy as printed at coulomb.F90:875).  Charges must be neutral
 * (canUseEwaldCode, coulomb.F90:255-265) and at most 1e4 in magnitude (EW_ERR_ARGUMENT: the
 * owned-block spreading accumulates each mesh point in fixed point, so an evaluation is bitwise
 * reproducible; see ew_options.tune and pair_storage for the shapes that are not).  energy may
 * be NULL. */
int ew_shift(ew_handle *h, const double *dq, double *shift, double *energy);

/* Same with device pointers (dq, shift: n doubles; energy: 1 double), enqueued on `stream`
 * (a hipStream_t; NULL = the handle's stream).  Asynchronous: no host synchronisation, no
 * neutrality or magnitude check (validate with ew_shift or on the host).  The handle's buffers
 * serve one evaluation at a time: an evaluation on another stream than the previous one is
 * ordered after it on the device, and every later call that rewrites or frees the handle's buffers
 * (ew_set_geometry, ew_gradients, ew_stress, ew_external_*, ew_destroy, evicting a captured
 * graph) waits for the last evaluation first.  The caller still owns its own buffers: d_dq
 * must stay valid and d_shift/d_energy unread until `stream` has passed the evaluation. */
int ew_shift_device(ew_handle *h, const double *d_dq, double *d_shift, double *d_energy,
                    void *stream);

/* ew_shift_device without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The evaluation's scratch (sorted charges, real-space
 * potentials, energy partials, mesh, spectrum and spreading copies) lives in one static device
 * array of 64 MiB per device, held until the process ends, instead of the handle's buffers; the
 * kernels are ew_shift_device's, launched directly (no graph is captured or instantiated).  A
 * geometry whose scratch exceeds 64 MiB gets EW_ERR_MEMORY (the 98,304-atom cell at 128^3 fits
 * in both precisions with direct spreading; FP64 at 192^3 does not).  The caller passes the
 * inputs and outputs as for ew_shift_device (d_energy may be NULL).  Every call on one device
 * shares that array, which keeps the last call's contents: calls that overlap on one device
 * race (order them on one stream or synchronise between them). */
int ew_shift_device_noalloc(ew_handle *h, const double *d_dq, double *d_shift, double *d_energy,
                            void *stream);

/* Several charge vectors through the current geometry in one grouped launch: the analogue of
 * nvec calls of TCoulomb%updateShifts on one structure (coulomb.F90:786-880), e.g. the charge
 * vectors of independent SCC cycles that share a geometry (spin or k-point replicas, trial
 * charges of a mixer).  The pair coefficients are read once for all vectors.  dq, shift:
 * [nvec][n] row-major; energy: nvec doubles (may be NULL).  1 <= nvec <= EW_MAX_VECTORS.  Each
 * vector's result is ew_shift's for that vector alone (to round-off: the grouped launch adds
 * the mesh and pair parts in one kernel).  Extra device memory per vector: about one mesh, one
 * spreading scratch
 * and four atom arrays, held until the next ew_set_geometry. */
#define EW_MAX_VECTORS 8
int ew_shift_multi(ew_handle *h, int nvec, const double *dq, double *shift, double *energy);

/* Same with device pointers (d_dq, d_shift: [nvec][n]; d_energy: nvec doubles or NULL),
 * enqueued on `stream` (NULL = the handle's stream); asynchronous, no neutrality check. */
int ew_shift_multi_device(ew_handle *h, int nvec, const double *d_dq, double *d_shift,
                          double *d_energy, void *stream);

/* Energy gradient of the Ewald part at charges dq, ADDED into grad[n][3] as +dU/dx: the
 * fftEwaldAddGradients convention (coulomb.F90:361-385, which subtracts the physical force
 * from DFTB+'s deriv array).  The short-range gamma gradient stays with TShortGamma. */
int ew_gradients(ew_handle *h, const double *dq, double *grad);

/* Device time per phase of the evaluations run since the last reset, in milliseconds, with
 * the number of evaluations; the analogue of the reference's per-phase NFFT timing
 * [spread, transform, deconvolve] (fe/nfft/nfft.f90:228-231, 255-303).  Needs opts.profile = 1
 * (the profiled evaluation synchronises its stream).  Phases, in order:
 *   0 charge gather   1 spread        2 FFT x (r2c)    3 FFT y          4 FFT z + influence
 *   5 inverse FFT y   6 FFT x (c2r)   7 pair sum       8 gather + combine + energy */
#define EW_NUM_PHASES 9
int ew_phase_times(ew_handle *h, double *ms, long long *calls, int reset);

/* Device time per span of the evaluations run since the last reset, in milliseconds, measured
 * by HIP events in the evaluation's own launch shape (opts.profile = 2; each evaluation waits
 * for its last event).  The spans on the evaluation stream are the critical path; the charge
 * gather and the pair sum run on the auxiliary stream beside it when the handle forks:
 *   0 charge gather (+ charge sum)   1 pair sum            2 spread: region kernel
 *   3 spread: assembly (or the whole spreading without privatised blocks)
 *   4 FFT x / fused x-y (r2c)        5 FFT y               6 FFT z + influence + z^-1
 *   7 inverse FFT y                  8 inverse FFT x / fused y-x (c2r)
 *   9 mesh gather (when the pair sum runs beside the chain the gather splits: the mesh
 *     potential at every atom before the join, the rest after it; else empty)
 *   10 wait for the pair sum         11 gather + combine + energy (split: combine + energy)
 *   12 whole evaluation (start to end on the evaluation stream)
 * No counterpart in the reference (its per-phase timing is ew_phase_times'). */
#define EW_NUM_TRACE 13
int ew_trace_times(ew_handle *h, double *ms, long long *calls, int reset);

/* TCoulomb%addStress (src/dftbp/dftb/coulomb.F90:721-757, invRStress 2777-2863): adds the Ewald
 * stress at charges dq into stress[3][3] (row-major, Hartree/Bohr^3): sigma = W/V with the
 * virial W = -dE/d(strain) -- SPME reciprocal virial, real-space pair virial within r_cut and,
 * for charged cells (opts.charged), the neutralising background.  Short gamma not included. */
int ew_stress(ew_handle *h, const double *dq, double *stress);

/* TCoulomb%getPotential, periodic branch (src/dftbp/dftb/coulomb.F90:3121-3160 ->
 * sumInvRPeriodicAsymm 1340-1418): the potential of m external point charges at the resident
 * atoms, potential[i] = sum_s q_s ewald(x_i - y_s) (caller order, overwritten), with ewald()
 * of coulomb.F90:2311-2345: Ewald reciprocal + real part, the background -pi/(V alpha^2) per
 * source (so a non-neutral set of charges is allowed), and -2 alpha/sqrt(pi) for a source on an
 * atom.  ext_coords[m][3] in Bohr (DFTB+'s chargeCoords(3, m)); blur_widths[m] (NULL: point
 * charges) gives Gaussian sources, erf(d/w)/d in place of 1/d within 10 w (ewaldReal,
 * 2523-2575); eps_soften > 0 replaces d by sqrt(d^2 + eps^2) in the real part (which, as in
 * the reference, makes the result depend on alpha: opts.alpha_rule = 1 gives DFTB+'s alpha).
 * The reciprocal part is the SPME convolution on the engine's mesh (its precision mode
 * applies).  A source within 1e-5 Bohr of an atom or of one of its periodic images counts as on
 * it (the reference tests the unwrapped difference only).  Uses the handle's mesh buffers: not
 * concurrent with other calls on the handle. */
int ew_external_potential(ew_handle *h, int m, const double *ext_coords,
                          const double *ext_charges, const double *blur_widths,
                          double eps_soften, double *potential);

/* TCoulomb%getPotentialGradient and %addExternalPotGrad, periodic branch
 * (src/dftbp/dftb/coulomb.F90:3162-3235 -> addInvRPrimePeriodicAsymm 1934-2092): derivatives of
 * the interaction of the resident atoms with m external point charges (same sources, widths
 * and conventions as ew_external_potential; no softening, as in the reference).
 *   dq != NULL (forces):  atom_grad[n][3] += dq_i sum_s q_s grad ewald(x_i - y_s) and, when
 *                         ext_grad != NULL, ext_grad[m][3] -= the same summed over the atoms;
 *   dq == NULL (tHamDeriv): atom_grad[n][3] += sum_s q_s grad ewald(x_i - y_s)  (ext_grad unused).
 * Both arrays are accumulated into, in DFTB+'s deriv(3, N) layout. */
int ew_external_gradients(ew_handle *h, int m, const double *ext_coords,
                          const double *ext_charges, const double *blur_widths,
                          const double *dq, double *atom_grad, double *ext_grad);

/* ---- one-shot calls, the reference library's own entry points ---- */

/* ewald_potential_force (fe/fft/ewald_fft_3d.f90:36-134): pot[i] = phi(x_i) and, when force
 * is not NULL, force[i][3] = -dU/dx_i (the physical force: SPME reciprocal force from the
 * B-spline derivatives plus the real-space pair force).  Ewald only (no short gamma).
 * opts may be NULL. */
int ew_potential_force(int n, const double *coords, const double *charges,
                       const double *latvecs, const ew_options *opts, double *pot,
                       double *force);

/* ewald_parameters (fe/fft/ewald_fft_3d.f90:142-191) for a configuration, without a device. */
int ew_parameters_for(int n, const double *charges, const double *latvecs,
                      const ew_options *opts, ew_parameters *out);

/* ewald_grid_memory (fe/fft/ewald_fft_3d.f90:197-244): predicted peak device memory of one
 * engine for this configuration, in gigabytes (1024^3 bytes); the admission predictor. */
int ew_grid_memory(int n, const double *charges, const double *latvecs, const ew_options *opts,
                   double *gigabytes);

/* ---- structure pools: the batched and serving forms ----
 *
 * The reference evaluates one structure per call (fe/fft/ewald_fft_3d.f90:36-134, called once
 * per SCC iteration from TCoulomb%updateShifts, coulomb.F90:786-880).  A pool keeps many
 * structures resident (geometry, pair lists, B-spline weights, FFT plans) and evaluates any
 * subset of them in one grouped launch per kernel; DFTB+ drivers that run many independent
 * SCC cycles (replicas, trajectories, screening) call it once per SCC iteration of the whole
 * set.  Each structure's result is the single-structure engine's to round-off, and bitwise the
 * same alone and in any batch. */
typedef struct ew_pool ew_pool;

/* capacity_bytes: device memory the pool may hold (0: no limit); ew_pool_add refuses a
 * structure that does not fit (EW_ERR_MEMORY), which is the admission rule.  Every structure
 * of a pool shares opts (precision, tol, r_cut, order); grids may differ.  Leave opts.tune at 0
 * (the fixed rules) for pools of many structures: the autotuner times each structure alone on
 * the whole device, and its choices can be slower in a launch shared with others. */
int ew_pool_create(const ew_options *opts, long long capacity_bytes, ew_pool **out);
int ew_pool_destroy(ew_pool *pool);
const char *ew_pool_last_error(const ew_pool *pool);

/* Adds one structure (arguments as ew_set_geometry; grid[3] overrides opts->grid, NULL keeps
 * it); *id receives its id, reused after ew_pool_remove. */
int ew_pool_add(ew_pool *pool, int n, const double *coords, const double *latvecs,
                const int *species, const double *hubbard_u, int n_species,
                const double *charges, const int *grid, int *id);
int ew_pool_remove(ew_pool *pool, int id);
int ew_pool_parameters(const ew_pool *pool, int id, ew_parameters *out);
/* Device bytes held by the pool's structures, its capacity and its structure count. */
int ew_pool_usage(const ew_pool *pool, long long *used, long long *capacity, int *count);

/* One SCC evaluation of `count` distinct structures: dq and shift are the structures' charge
 * and shift vectors concatenated in `ids` order (caller atom order within each), energies[i]
 * the energy of structure ids[i]. */
int ew_pool_shift(ew_pool *pool, int count, const int *ids, const double *dq, double *shift,
                  double *energies);
/* The same with device buffers (host arrays of device pointers, one per structure), enqueued
 * on `stream` (a hipStream_t; NULL = the pool's stream) without synchronising.  Evaluations of
 * one structure must not overlap in time; a structure may appear once per call. */
int ew_pool_shift_device(ew_pool *pool, int count, const int *ids, const double *const *d_dq,
                         double *const *d_shift, double *d_energies, void *stream);
/* Waits for the pool's stream. */
int ew_pool_sync(ew_pool *pool);

/* ---- serving: a request loop with continuous batching over a pool ----
 *
 * Requests are SCC charge updates for resident structures, submitted from any thread.  A worker
 * launches a batch whenever a slot is free (at most max_inflight batches on the GPU at once, each
 * on its own stream): every pending request whose structure has nothing in flight, in arrival
 * order, up to max_batch structures in one grouped evaluation.  While a server runs on a pool,
 * only the server may evaluate or change the pool.  No reference counterpart: each request is
 * one ewald_potential_force call (fe/fft/ewald_fft_3d.f90:36-134) on a resident structure. */
typedef struct ew_server ew_server;

typedef struct {
  int max_batch;    /* structures per grouped launch; 0: 64                                  */
  int max_inflight; /* batches on the GPU at once; 0: 2                                      */
  int reserved[6];
} ew_server_options;

typedef struct {
  long long submitted, completed, failed, batches;
  double mean_batch;          /* requests per batch                                      */
  double p50_ms, p99_ms, max_ms; /* latency from submission to the result in the caller's buffers */
} ew_server_report;

int ew_server_create(ew_pool *pool, const ew_server_options *opts, ew_server **out);
/* Finishes the requests already submitted, then stops. */
int ew_server_destroy(ew_server *server);
const char *ew_server_last_error(const ew_server *server);
/* Queues one evaluation of structure id: dq is copied at submission; shift[n] (and *energy
 * when not NULL) are written when the request completes and must stay valid until then. */
int ew_server_submit(ew_server *server, int id, const double *dq, double *shift, double *energy,
                     long long *ticket);
/* Waits for a ticket (timeout_s < 0: no limit); EW_OK, the request's error, or
 * EW_ERR_PENDING at the timeout (the ticket stays valid). */
int ew_server_wait(ew_server *server, long long ticket, double timeout_s);
/* Waits until nothing is pending or in flight. */
int ew_server_drain(ew_server *server);
/* Counters and latency percentiles since the last reset. */
int ew_server_stats(ew_server *server, ew_server_report *out, int reset);

/* ---- SCC charge mixing on the device ----
 *
 * DFTB+'s charge mixers (src/dftbp/mixer/: the modified Broyden mixer TBroydenMixerReal,
 * broydenmixer.F90:114-342, and the simple mixer, simplemixer.F90:77-92) with their history
 * resident on the device, so that an SCC cycle keeps the charges on the GPU between
 * ew_shift_device and the mixer (the call in TScc / the SCC loop that mixes qInpRed with
 * qDiffRed).  A Broyden mix is three kernels on the given stream and no host round trip. */
typedef struct ew_mixer ew_mixer;

#define EW_MIXER_BROYDEN 0    /* modified Broyden (Johnson, PRB 38, 12807 (1988)), the default */
#define EW_MIXER_SIMPLE 1     /* qInpResult + mixParam * qDiff                                  */

/* Mirrors TBroydenMixerInp (broydenmixer.F90:31-50); zero fields take its defaults. */
typedef struct ew_mixer_options {
  int kind;            /* EW_MIXER_BROYDEN or EW_MIXER_SIMPLE                                 */
  int max_iter;        /* maxSccIter: mixes per reset (Broyden keeps max_iter - 1 vectors of
                          history); 0 -> 100, at most 1024                                      */
  double mix_param;    /* mixParam; 0 -> 0.2                                                    */
  double omega0;       /* weight of the Jacobian differences; 0 -> 0.01                        */
  double min_weight;   /* 0 -> 1.0                                                              */
  double max_weight;   /* 0 -> 1e5                                                              */
  double weight_fac;   /* numerator of the weights; 0 -> 1e-2                                   */
  int device;          /* HIP device ordinal                                                    */
  int reserved[3];
} ew_mixer_options;

int ew_mixer_create(const ew_mixer_options *opts, ew_mixer **out);
int ew_mixer_destroy(ew_mixer *mixer);
const char *ew_mixer_last_error(const ew_mixer *mixer);
/* TMixer%reset (broydenmixer.F90:144-169): vectors of n_elem entries, history cleared. */
int ew_mixer_reset(ew_mixer *mixer, int n_elem);
/* TMixer%mix (broydenmixer.F90:180-200): q_inp_result holds the input charges on entry and the
 * mixed charges on exit; q_diff = output - input charges.  More than max_iter mixes since the
 * reset fail with EW_ERR_ARGUMENT, as the reference stops ("Maximal nr. of steps exceeded"). */
int ew_mixer_mix(ew_mixer *mixer, double *q_inp_result, const double *q_diff);
/* Same on device arrays of n_elem doubles, enqueued on `stream` (NULL = the mixer's own
 * stream, which like the engine's is ordered with the null stream); asynchronous.  The calls
 * on one mixer must be ordered (one stream, or synchronised). */
int ew_mixer_mix_device(ew_mixer *mixer, double *d_q_inp_result, const double *d_q_diff,
                        void *stream);
/* Mixes since the last reset. */
int ew_mixer_iteration(const ew_mixer *mixer, int *iter);

#ifdef __cplusplus
}
#endif
#endif
