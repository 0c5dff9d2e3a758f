/* C-ABI of libdgfs_hip.so: the MI355X (gfx950) right-hand-side hot path of
 * the DGFS Boltzmann solver.
 *
 * Every entry point returns 0 on success or a nonzero dgfs_status; the text of
 * the last failure on the calling thread is available from dgfs_last_error().
 * The mapping of status to exception follows the reference's ctypes wrappers
 * (frfs/solvers/dgfs/cufft.py:108-117, frfs/backends/cuda/cublas.py:75-80).
 *
 * State buffers (distribution registers, operator matrices, geometry) are
 * caller-owned device pointers, as in the reference where the backend
 * allocates and kernels receive raw pointers (frfs/backends/cuda/base.py:93-102).
 * Plans/handles are library-owned.  All work is stream ordered on the stream
 * argument (a hipStream_t passed as void*; NULL = the null stream).
 */
#ifndef DGFS_HIP_H
#define DGFS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum dgfs_status {
  DGFS_OK = 0,
  DGFS_ERR_INVALID_VALUE = 1,   /* bad argument or shape */
  DGFS_ERR_UNSUPPORTED = 2,     /* size or option this build does not support */
  DGFS_ERR_ALLOC = 3,           /* device allocation failed */
  DGFS_ERR_HIP = 4,             /* a HIP runtime call failed */
  DGFS_ERR_NOT_INITIALIZED = 5  /* null or destroyed handle */
};

/* Text of the last error raised on this thread ("" if none). */
const char* dgfs_last_error(void);

/* Library version string. */
const char* dgfs_version(void);

/* ------------------------------------------------------------------------
 * Fast-spectral VHS collision operator Q(f, f).
 *
 * Replaces the reference's DGFSVHSGLLScatteringModel
 * (frfs/solvers/dgfs/scattering.py:44-286): perform_precomputation (:67-224,
 * tables of vhs-gll.mako:45-101, cuFFT plans :141-150) becomes
 * dgfs_vhs_create; fs(d_arr_in, d_arr_out, elem, upt) (:227-286) becomes
 * dgfs_vhs_eval over a range of elements and all solution points in one call.
 * ---------------------------------------------------------------------- */
typedef struct dgfs_vhs_desc {
  int Nv;           /* velocity points per axis (Nv^3 = vsize); 2^a 3^b, even */
  int Nrho;         /* radial Gauss points (velocitymesh.py:48) */
  int M;            /* half-sphere design points (velocitymesh.py:130) */
  int precision;    /* 64: FP64 transforms; 32: FP32 transforms */
  int project;      /* 1: project Q onto the 5 collision invariants in FP64 */
  double L;         /* half-width of the velocity box (velocitymesh.py:62) */
  double gamma;     /* 2 (1 - omega) (scattering.py:53) */
  double prefac;    /* kernel prefactor (scattering.py:61-62) */
  double sw;        /* spherical weight 2 pi / M (velocitymesh.py:135) */
  double cw;        /* velocity cell weight (2L/Nv)^3 (velocitymesh.py:72) */
  const double* qz; /* [Nrho] radial nodes on [0, R] (velocitymesh.py:125) */
  const double* qw; /* [Nrho] radial weights (velocitymesh.py:126) */
  const double* sz; /* [M][3] half-sphere directions (velocitymesh.py:134) */
} dgfs_vhs_desc;

/* Precompute the spectral weights b1, b2 and size the device workspace.
 * Replaces DGFSVHSGLLScatteringModel.perform_precomputation
 * (frfs/solvers/dgfs/scattering.py:67-224). */
int dgfs_vhs_create(const dgfs_vhs_desc* desc, void** handle);

/* Q at every solution point of elements [e0, e0+ne) of one element type.
 * f, Q: the full distribution registers in the reference's AoSoA layout,
 * element (upt, r, elem) at upt*ldim + ((elem/32)*vsize + r)*32 + elem%32
 * (vhs-gll.mako:222-223, backends/base/types.py:21-43).  f and Q may not alias.
 * Replaces the per-point loop of DGFSSystem.collide (frfs/solvers/dgfs/
 * system.py:199-218) over DGFSVHSGLLScatteringModel.fs (scattering.py:227-286). */
int dgfs_vhs_eval(void* handle, const double* f, double* Q, int nupts, int ldim,
                  int vsize, int neles, int e0, int ne, void* stream);

/* Q for npts independent points stored contiguously, f[pt*vsize + r].
 * Spatially homogeneous runs (a uniform state per point) and the batch
 * benchmark use this form; it has no reference counterpart beyond fs(). */
int dgfs_vhs_eval_contig(void* handle, const double* f, double* Q, int npts, void* stream);

/* dgfs_vhs_eval_contig that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot call hipMalloc (code under stream
 * capture, tight serving loops, harnesses that forbid allocation).  The
 * workspace is a 256 MiB device array inside the library, one per GPU, held
 * until the process ends; points are evaluated in chunks that fit it, and a
 * configuration whose single point does not fit returns DGFS_ERR_UNSUPPORTED.
 * Every call on one GPU shares that GPU's array (and the handle's launch
 * state), so calls that overlap on one GPU race.  Per-kernel timing is not
 * recorded.  Q agrees with dgfs_vhs_eval_contig up to the summation order of
 * the radial groups. */
int dgfs_vhs_eval_contig_noalloc(void* handle, const double* f, double* Q, int npts, void* stream);

/* Two independent contiguous batches on two velocity grids (32+24, 32+16 or
 * 24+16 points per axis, same precision) evaluated together.  With the
 * multi-pass gain (FP64, or when selected) the gain work items of both grids
 * share one persistent launch.  When both handles take a register-resident
 * gain (FP32), whose kernels need different workgroup sizes, the two
 * pipelines run back to back on the stream, each gain grid filling the chip;
 * Q then equals two separate calls bitwise.  Batches beyond one workspace
 * chunk run as separate launches.  No reference counterpart (the reference
 * has one velocity grid per run). */
int dgfs_vhs_eval_pair(void* handle_a, const double* f_a, double* Q_a, int npts_a, void* handle_b,
                       const double* f_b, double* Q_b, int npts_b, void* stream);

/* Release the handle's device memory (cufft plan destroy analogue). */
int dgfs_vhs_destroy(void* handle);

/* Work split chosen for the last eval: points per launch, radial groups per
 * point, persistent workgroups.  Used by the autotuner and the benchmark. */
int dgfs_vhs_get_config(void* handle, int* ngroups, int* nslots, int* threads);
/* The grid a handle was built for: Nv (points per axis), Nrho, M,
 * precision (64 or 32) and the half-width L of the velocity box.  Lets a
 * caller that shares a handle (the surrogate's fallback) check shapes before
 * any launch; no reference counterpart (the reference holds one mesh). */
int dgfs_vhs_get_dims(void* handle, int* Nv, int* Nrho, int* M, int* precision, double* L);
/* Override the radial-group split (0 = automatic). */
int dgfs_vhs_set_groups(void* handle, int ngroups);
/* Cap the persistent grid of the multi-pass kernels at nslots workgroups
 * (0 = the occupancy limit).  With set_groups and set_algorithm, the launch
 * shape the autotuner (dgfs_hip/autotune.py) searches. */
int dgfs_vhs_set_slots(void* handle, int nslots);

/* Gain-kernel algorithm: 0 automatic, 1 multi-pass (slab + pencil passes
 * through a per-workgroup scratch, any N and precision), 2 register-resident
 * (the whole N^3 field in registers, FP32 with Nv 16 or 32; Nyquist-plane
 * corrections precomputed by a separate kernel), 3 register-resident with the
 * Nyquist-plane corrections computed inside the gain kernel, 4 register-resident
 * with one pencil per thread and the transposes through LDS (FP32 with Nv 16
 * or 24; Nyquist planes from the separate kernel), 5 as 2 with the in-plane
 * transposes wave-local (no workgroup barrier between the z, y passes), 6 slab
 * kernel (FP64 or FP32, Nv 16, 24 or 32: a workgroup holds N/NS x-planes of
 * the inverse transforms in LDS, NS = 1, 3, 4 slabs per point, with the
 * x-transform decimated to the slab's planes; a second kernel completes the
 * forward transform along x and sums over the radial nodes).  The automatic
 * choice is, in FP32, 4 at Nv 16 and 24 and 5 at Nv 32; in FP64, 6 at Nv 16,
 * 24 and 32; 1 otherwise.
 * get_algorithm reports the one the last eval ran.
 * No reference counterpart (the reference has one cuFFT dataflow). */
int dgfs_vhs_set_algorithm(void* handle, int algorithm);
int dgfs_vhs_get_algorithm(void* handle, int* algorithm);

/* Kernel timing with HIP events recorded on the launch stream around every
 * launch (no added synchronisation).  set_timing(on) also clears the totals;
 * get_timing synchronises on the recorded events and returns, for the kernel
 * classes {prep, gain, finalize, project, nyquist}, the summed milliseconds
 * and the number of launches.  The reference only prints wall-clock totals
 * (integrators/base.py:265-279); this has no direct counterpart. */
int dgfs_vhs_set_timing(void* handle, int on);
int dgfs_vhs_get_timing(void* handle, double* ms /*[5]*/, int* counts /*[5]*/);

/* ------------------------------------------------------------------------
 * Multi-species fast-spectral VSS collision operator
 * Q_p = sum_q Q_pq(f_p, f_q), up to 4 species.
 *
 * Replaces the reference's DGFSBiVSSGLLScatteringModel
 * (frfs/solvers/dgfsbi/scattering.py:46-330): perform_precomputation
 * (:96-257, tables of vss-gll.mako:31-95) becomes dgfs_vss_create;
 * fs(elem, upt, *args) (:323-330), i.e. Qij (:259-320) over every ordered
 * pair, becomes dgfs_vss_eval over a range of elements.
 * ---------------------------------------------------------------------- */
typedef struct dgfs_vss_desc {
  int Nv;             /* velocity points per axis */
  int Nrho;           /* radial Gauss points (dgfsbi/velocitymesh.py:47) */
  int M;              /* full-sphere design points (dgfsbi/velocitymesh.py:131-137) */
  int Mpre;           /* design points of the b/c precomputation (scattering.py:53-56) */
  int nspcs;          /* species, 1..4 */
  int precision;      /* 64 or 32 */
  int project;        /* 1: project onto species masses, total momentum, total energy */
  double L;           /* half-width of the velocity box */
  double sw;          /* 4 pi / M */
  double cw;          /* (2L/Nv)^3 */
  const double* qz;   /* [Nrho] radial nodes on [0, R] */
  const double* qw;   /* [Nrho] radial weights */
  const double* sz;   /* [M][3] full-sphere directions */
  const double* szpre;/* [Mpre][3] precomputation directions */
  const double* masses; /* [nspcs] nondimensional molecular masses */
  const double* gamma;  /* [nspcs][nspcs] 2 (1 - omega_pq) (scattering.py:66-67) */
  const double* eta;    /* [nspcs][nspcs] alpha_pq - 1 (scattering.py:63-64) */
  const double* prefac; /* [nspcs][nspcs] (scattering.py:79-86) */
} dgfs_vss_desc;

/* Build b_pq (Nrho*M*Nv^3 complex per pair) and c_pq on the device.
 * Replaces DGFSBiVSSGLLScatteringModel.perform_precomputation
 * (frfs/solvers/dgfsbi/scattering.py:96-257). */
int dgfs_vss_create(const dgfs_vss_desc* desc, void** handle);

/* Q of every species at every solution point of elements [e0, e0+ne).
 * f[s], Q[s]: host arrays of nspcs device pointers to the species' AoSoA
 * registers (same layout as dgfs_vhs_eval); Q[p] is overwritten with
 * sum_q Q_pq.  Replaces the per-point loop over DGFSBiVSSGLLScatteringModel.fs
 * (frfs/solvers/dgfsbi/scattering.py:323-330, called from
 * frfs/solvers/dgfsbi/system.py:239-251). */
int dgfs_vss_eval(void* handle, const double* const* f, double* const* Q, int nupts, int ldim,
                  int vsize, int neles, int e0, int ne, void* stream);

/* Same for npts points stored contiguously per species, f[s][pt*vsize + r]. */
int dgfs_vss_eval_contig(void* handle, const double* const* f, double* const* Q, int npts,
                         void* stream);
int dgfs_vss_destroy(void* handle);
int dgfs_vss_get_config(void* handle, int* ngroups, int* nslots, int* threads);
int dgfs_vss_set_groups(void* handle, int ngroups);
/* Cap the persistent grid (and so the scratch cubes in use) at nslots
 * workgroups; 0 = the occupancy limit. */
int dgfs_vss_set_slots(void* handle, int nslots);
/* Gain kernel: 0 automatic (the slab kernel with isotropic scattering at Nv
 * 16 and 24, either precision; register-resident in FP32 at Nv 16 and 24
 * otherwise; multi-pass otherwise), 1 multi-pass (any precision and grid), 4 register-resident with
 * one pencil per thread (FP32 at Nv 16 and 24), 6 its isotropic form with t3
 * and t4 from one packed complex transform plus Nyquist-plane corrections (not
 * the automatic choice: slower than 4 as measured), 7 slab kernel (eta = 0 for
 * every pair, Nv 16 or 24, either precision: x-slabs of t3 and t4 in LDS, a
 * second kernel for the x-forward transform and the sum over radial nodes).
 * With eta = 0 for every pair (isotropic scattering) 1, 4 and 7 sum the
 * direction products in space before one forward transform per radial node.
 * get_algorithm reports the kernel of the last call: 1, 4, 5 for the isotropic
 * register-resident kernel, 6 or 7. */
int dgfs_vss_set_algorithm(void* handle, int algorithm);
int dgfs_vss_get_algorithm(void* handle, int* algorithm);

/* Antipodal sharing of the gain products (on by default when the direction
 * design holds the exact negation of every direction): one product per
 * (radial node, direction) serves both ordered pairs (p, q) and (q, p), and a
 * self pair runs over half of the directions with the two antipodal table
 * rows summed.  The result equals the unshared sum up to rounding; on = 0
 * selects the reference's per-pair loop.  Refines 'gainsum' of
 * frfs/solvers/dgfsbi/scattering.py:259-320 (no counterpart switch there). */
int dgfs_vss_set_symmetry(void* handle, int on);
int dgfs_vss_get_symmetry(void* handle, int* on);
/* Kernel classes {prep, gain, finalize, project, unused}. */
int dgfs_vss_set_timing(void* handle, int on);
int dgfs_vss_get_timing(void* handle, double* ms /*[5]*/, int* counts /*[5]*/);

/* ------------------------------------------------------------------------
 * HIP graphs of time steps.  The reference launches every kernel of a step
 * from Python (integrators/std/steppers.py:122-225); here a step is captured
 * once on the solver's (non-default) stream and replayed with one launch.
 * No reference counterpart.
 * ---------------------------------------------------------------------- */
int dgfs_graph_begin(void* stream);
int dgfs_graph_end(void* stream, void** graph);
int dgfs_graph_launch(void* graph, void* stream);
int dgfs_graph_nodes(void* graph, long* count);
int dgfs_graph_destroy(void* graph);

/* ------------------------------------------------------------------------
 * Learned collision surrogate (no reference counterpart: the reference has
 * only the spectral operator).  A point's distribution is encoded by its 8^3
 * Hermite modes in the frame of its own Maxwellian; a 4-layer spectral
 * network (width 32, GEMMs on MFMA, BF16 operands, FP32 accumulation) maps
 * them to the modes of Q scaled by prefac rho T^(gamma/2); the decoded Q is
 * projected onto the collision invariants in FP64.  Points whose
 * non-equilibrium indicator ||f - M[f]|| / ||f|| exceeds a threshold fall back
 * to the exact operator.  Parameters are one FP32 vector of
 * dgfs_sur_param_count() values owned by the caller:
 *   P0[32], P1[32], E[512][32],
 *   4 x { K[512][512], W[32][32], R[32][32], bias[32] },
 *   Pout[32], bout[512].
 * ---------------------------------------------------------------------- */
typedef struct dgfs_sur_desc {
  int Nv;          /* velocity points per axis: 8, 12, 16, 24 or 32 */
  int max_batch;   /* largest network batch (activation workspace) */
  double L;        /* half-width of the velocity box */
  double cw;       /* (2L/Nv)^3 */
  double gamma;    /* 2 (1 - omega) */
  double prefac;   /* collision kernel prefactor */
} dgfs_sur_desc;

long dgfs_sur_param_count(void);
int dgfs_sur_create(const dgfs_sur_desc* desc, void** handle);
int dgfs_sur_destroy(void* handle);
/* out[P][512] = network(x[P][512], lt[P] = log T); bf16 = 1: BF16 MFMA
 * operands, 0: FP32 MFMA operands.  Keeps the activations for backward. */
int dgfs_sur_forward(void* handle, const float* params, const float* x, const float* lt, int P,
                     float* out, int bf16, void* stream);
/* grad = d(loss)/d(params) given dout = d(loss)/d(out) of the last forward. */
int dgfs_sur_backward(void* handle, const float* params, const float* dout, float* grad, void* stream);
/* One Adam step (bias-corrected, step >= 1) over n parameters. */
int dgfs_sur_adam(long n, float* params, const float* grad, float* m1, float* m2, float lr, float b1,
                  float b2, float eps, int step, void* stream);
/* f[npts][Nv^3] -> mom[npts][8] (rho, ux, uy, uz, T, indicator ||f - M[f]|| / ||f||,
 * high-mode indicator = share of f's norm in modes of total degree >= 4, 0),
 * x[npts][512] Hermite modes (a_000 stored as 0), lt[npts] = log T; x, lt may be null. */
int dgfs_sur_encode(void* handle, const double* f, int npts, double* mom, float* x, float* lt,
                    void* stream);
/* Training targets: Hermite modes of Q[npts][Nv^3] in the frames of mom,
 * divided by prefac rho T^(gamma/2). */
int dgfs_sur_targets(void* handle, const double* Q, const double* mom, int npts, float* y,
                     void* stream);
/* Q[idx[i]] = decoded network output row i (idx null: row i -> point i);
 * project = 1 (idx null only) applies the FP64 invariant projection. */
int dgfs_sur_decode(void* handle, const float* y, const double* mom, const int* idx, int nidx,
                    double* Q, int project, void* stream);
/* Surrogate with fallback to the exact operator `vhs` (a dgfs_vhs handle of
 * the same grid) for points whose indicator exceeds `threshold` or (when
 * threshold_hi > 0) whose high-mode indicator exceeds `threshold_hi`;
 * synchronises the stream once to read the indicators. */
int dgfs_sur_eval(void* handle, void* vhs, const float* params, const double* f, double* Q, int npts,
                  double threshold, double threshold_hi, int bf16, int* nfallback, void* stream);

/* Settings of the sampled spectral check. */
typedef struct dgfs_sur_check {
  double tol;          /* target relative L2 error of the batch's Q */
  double safety;       /* the kept estimate must be within (safety * tol)^2 ||Q||^2; 0 < safety <= 1 */
  double sample_frac;  /* share of each bin's candidates evaluated exactly as the sample */
  int nbins;           /* indicator-ranked bins of the surrogate candidates */
  int min_per_bin;     /* smallest sample of a bin */
  unsigned long long seed;  /* offset of the stratified sample */
} dgfs_sur_check;

typedef struct dgfs_sur_check_stats {
  int nfallback;       /* points given the exact operator for cause (indicator or rejected bin) */
  int nsampled;        /* points of the sample (exact Q kept) */
  int nsurrogate;      /* points whose Q is the network's */
  int nrejected_bins;  /* bins sent to the exact operator by the check */
  double est_rel_err;  /* estimated relative L2 error of the kept surrogate points */
} dgfs_sur_check_stats;

/* dgfs_sur_eval with a sampled spectral check.  The surrogate candidates of
 * the indicator routing are ranked by indicator and cut into nbins bins of
 * equal count; a stratified sample of every bin is evaluated by the network
 * and by the exact operator.  Per bin, the sample's squared errors and |Q|^2
 * per unit squared indicator (a ratio estimator) predict the unsampled
 * points' from their indicators; whole bins, worst predicted error per point
 * first, fall back to the exact operator until the predicted error of the
 * points the network keeps is within safety * tol of ||Q||.  Sampled points
 * keep their exact Q.  Synchronises the stream twice (indicators, sample). */
int dgfs_sur_eval_checked(void* handle, void* vhs, const float* params, const double* f, double* Q,
                          int npts, double threshold, double threshold_hi, const dgfs_sur_check* check,
                          int bf16, dgfs_sur_check_stats* stats, void* stream);
/* dst[i] = src[i] rounded to BF16 (nearest even), stored as FP32: the BF16
 * weights the network runs with in its FP32-operand mode. */
int dgfs_sur_round_bf16(long n, const float* src, float* dst, void* stream);
/* C = alpha op(A) op(B) + beta C (row-major FP32 storage; bf16 as above);
 * the network's GEMM, exported for its tests. */
int dgfs_nn_gemm(int transA, int transB, int M, int N, int K, float alpha, const float* A, long lda,
                 const float* B, long ldb, float beta, float* C, long ldc, int bf16, void* stream);

/* ------------------------------------------------------------------------
 * Flux-reconstruction advection of the velocity columns and RK stage kernels.
 * Matrices are the reference's AoSoA layout: element (row, var v, elem e) at
 * row*ld + ((e/32)*nvars + v)*32 + e%32; velocities are the variables.
 * ---------------------------------------------------------------------- */

/* C = alpha A B + beta C with a constant operator A [m][k] (row stride lda,
 * k <= 64), B [k][ldb], C [m][ldc], ncol columns; FP64 on MFMA.
 * Replaces backend.kernel('mul', a, b, out, alpha, beta):
 * CUDACUBLASKernels.mul (frfs/backends/cuda/cublas.py:105-137) and
 * CUDAGiMMiKKernels.mul (frfs/backends/cuda/gimmik.py:18-48). */
int frfs_mul(const double* A, int m, int k, long lda, const double* B, long ldb, double* C,
             long ldc, long ncol, double alpha, double beta, void* stream);

/* Transformed advective flux f[i][upt][v][e] = sum_k smats[i][upt][k][e]
 * c_k[vidx+v] u[upt][v][e]; cv is [3][vsize].  Replaces the 'tflux' pointwise
 * kernel (frfs/solvers/dgfs/kernels/tflux.mako:6-25, flux.mako:4-20). */
int dgfs_tflux(int ndims, int nupts, int neles, int nvars, int vidx, const double* u, long ldu,
               const double* smats, long ldsm, double* f, long ldf, const double* cv, int vsize,
               void* stream);

/* out = -rcpdjac out (no source term).  Replaces 'negdivconf'
 * (frfs/solvers/dgfs/kernels/negdivconf.mako:5-14). */
int dgfs_negdivconf(int nupts, int neles, int nvars, double* out, long ld, const double* rcpdjac,
                    long ldr, void* stream);

/* Interior common flux: exact upwind fn = 1/2 n.c (ul+ur) + 1/2 |n.c| (ul-ur),
 * then ul = |n| fn, ur = -|n| fn.  Interface point i of the left/right side
 * is at L + loff[i] (+ 32 v for velocity v); nl is [ndims][nfpts].  Replaces
 * 'intcflux' with the rusanov solver (frfs/solvers/dgfs/kernels/
 * intcflux.mako:7-25, rsolvers/rusanov.mako:5-52; views of
 * frfs/solvers/dgfs/inters.py:17-58). */
int dgfs_intcflux(int ndims, int nfpts, int nvars, int vidx, double* L, const long* loff, double* R,
                  const long* roff, const double* nl, const double* magnl, const double* cv,
                  int vsize, void* stream);

/* Boundary common flux, bctype 0 diffuse wall, 1 specular wall, 2 inlet,
 * 3 outlet, 4 vanish, 5 normal-shock inlet, 6 diffuse wall with per-point
 * wall velocity and temperature: ghost state then the upwind flux,
 * ul = |n| fn.  f0: wall/inlet Maxwellian [vsize], or for bctype 6 the wall
 * state [4][nfpts] (Ux, Uy, Uz, T, nondimensional) whose unnormalised
 * Maxwellian exp(-|c-U|^2/T) is formed in the kernel; bc_vals: diffuse
 * [2][nfpts] flux balance, specular stored states [vsize][nfpts]; rmap:
 * specular reflection map [vsize].  Replaces 'bccflux' (frfs/solvers/dgfs/
 * kernels/bccflux.mako:8-39 with bcs/dgfs-wall-diffuse.mako,
 * dgfs-wall-specular.mako, dgfs-inlet.mako, dgfs-outlet.mako,
 * dgfs-vanish.mako, dgfs-inlet-normalshock.mako, dgfs-wall-diffuse-cyl.mako;
 * frfs/solvers/dgfs/inters.py:175-492). */
int dgfs_bccflux(int bctype, int ndims, int nfpts, int nvars, int vidx, double* L, const long* loff,
                 const double* nl, const double* magnl, const double* f0, const double* bc_vals,
                 const int* rmap, double Ux, double Uy, double Uz, const double* cv, int vsize,
                 void* stream);

/* Boundary update over one velocity batch: diffuse wall accumulates
 * bc0 += sum_{(c-U).n>=0} (c-U).n ul cw, bc1 += sum_{<0} (c-U).n f0 cw (reset
 * when vidx == 0); bctype 6 does the same with the per-point wall state of
 * dgfs_bccflux; specular stores ul.  Replaces 'bccupdate'
 * (frfs/solvers/dgfs/kernels/bccupdate.mako:8-24, bcs/dgfs-wall-diffuse-
 * update.mako:4-40, dgfs-wall-diffuse-cyl-update.mako:4-48,
 * dgfs-wall-specular-update.mako:4-16). */
int dgfs_bccupdate(int bctype, int ndims, int nfpts, int nvars, int vidx, const double* L,
                   const long* loff, const double* nl, const double* f0, double* bc_vals, double Ux,
                   double Uy, double Uz, double cw, const double* cv, int vsize, void* stream);

/* x[0] = a[0] x[0] + sum_{i>0} a[i] x[i] over n values, nv <= 8 (x, a are host
 * arrays of device pointers / coefficients).  Replaces 'axnpby'
 * (frfs/backends/cuda/blasext.py:14-43, kernels/axnpby.mako:5-36). */
int frfs_axnpby(long n, int nv, double* const* x, const double* a, void* stream);

/* Velocity-slice copy dst[.., d0+v, ..] = src[.., s0+v, ..] for v < count
 * between registers of nvsrc and nvdst variables.  Replaces 'copy_to_reg' /
 * 'copy_from_reg' (frfs/backends/cuda/blasext.py:77-138,
 * kernels/copy_dgfs.mako:5-26). */
int dgfs_copy_reg(int nupts, int neles, const double* src, long lds, int nvsrc, int s0,
                  double* dst, long ldd, int nvdst, int d0, int count, void* stream);

/* Macroscopic moments of the full register f (nupts, vsize, neles) into
 * out[upt][m][e], m over rho, Ux, Uy, Uz, T, qx, qy, qz, Pxx, Pxy, Pxz, Pyy,
 * Pyz, Pzz, p.  Replaces the host-side DGFSMomWriterStdPlugin._compute_moments
 * (frfs/plugins/dgfsmomwriterstd.py:14-214). */
int dgfs_moments(int nupts, int neles, int vsize, const double* f, long ld, const double* cv,
                 double cw, double* out, void* stream);

/* Residual monitor (frfs/plugins/dgfsresidualstd.py:41-49): out2[0] =
 * sum (curr - prev)^2, out2[1] = sum prev^2 over the valid entries of two
 * full-velocity AoSoA registers of leading dimension ld (row and AoSoA
 * padding excluded); work is caller-owned device scratch of 1024 doubles.
 * The host takes sqrt(out2[0] / out2[1]) as the residual. */
int dgfs_residual_norms(int nupts, int neles, int vsize, const double* prev, const double* curr, long ld,
                        double* out2, double* work, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DGFS_HIP_H */
