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

/* Release the handle's device memory (cufft plan destroy analogue). */
int dgfs_vhs_destroy(void* handle);

/* Work split chosen for the last eval: points per launch, radial groups per
 * point, persistent workgroups.  Used by the autotuner and the benchmark. */
int dgfs_vhs_get_config(void* handle, int* ngroups, int* nslots, int* threads);
/* Override the radial-group split (0 = automatic). */
int dgfs_vhs_set_groups(void* handle, int ngroups);

/* Kernel timing with HIP events recorded on the launch stream around every
 * launch (no added synchronisation).  set_timing(on) also clears the totals;
 * get_timing synchronises on the recorded events and returns, for the kernel
 * classes {prep, gain, finalize, project}, the summed milliseconds and the
 * number of launches.  The reference only prints wall-clock totals
 * (integrators/base.py:265-279); this has no direct counterpart. */
int dgfs_vhs_set_timing(void* handle, int on);
int dgfs_vhs_get_timing(void* handle, double* ms /*[4]*/, int* counts /*[4]*/);

#ifdef __cplusplus
}
#endif

#endif /* DGFS_HIP_H */
