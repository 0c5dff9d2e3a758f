// Flux-reconstruction advection kernels of the DGFS right-hand side for gfx950.
//
// Every kernel works on the reference's AoSoA matrices: element (row, var v,
// elem e) of a matrix with nv variables lives at row*ld + ((e/32)*nv + v)*32
// + e%32 (frfs/backends/base/types.py:21-43).  Velocities are the variables,
// so a kernel loops over the velocity index instead of unrolling it at
// code-generation time as the reference's Mako templates do (tflux.mako:20,
// rusanov.mako:30); threads run along the contiguous 32-element AoSoA blocks
// so HBM accesses coalesce.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "common.hpp"

namespace dgfs {

// ---------------------------------------------------------------------------
// operator GEMM on MFMA: C = alpha A B + beta C
// A [m][k] row-major constant operator (m, k <= 64), B [k][ldb], C [m][ldc].
// One wave computes 16 x 16 output tiles with v_mfma_f64_16x16x4f64:
//   A operand: lane l holds A[l%16][4s + l/16]
//   B operand: lane l holds B[4s + l/16][col0 + l%16]
//   D result : lane l holds D[4*r + l/16][l%16], r = 0..3 (rows interleaved
//              over the four 16-lane groups)
// A is preloaded into registers once per wave; the wave strides over column
// tiles, four tiles per iteration for memory-level parallelism.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kMulMaxK = 64;
constexpr int kMulTiles = 4;

__global__ __launch_bounds__(256) void k_mul_mfma(const double* __restrict__ A, int m, int k, long lda,
                                                  const double* __restrict__ B, long ldb,
                                                  double* __restrict__ C, long ldc, long ncol,
                                                  double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int ksteps = (k + 3) / 4;
  const int rtiles = (m + 15) / 16;
  const long ctiles = (ncol + 15) / 16;
  const long ngroups = (ctiles + kMulTiles - 1) / kMulTiles;
  const long wave_id = static_cast<long>(blockIdx.x) * (blockDim.x >> 6) + wave;
  const long nwaves = static_cast<long>(gridDim.x) * (blockDim.x >> 6);

  for (int rt = 0; rt < rtiles; ++rt) {
    double a[kMulMaxK / 4];
    const int row = rt * 16 + li;
#pragma unroll
    for (int s = 0; s < kMulMaxK / 4; ++s) {
      const int kk = 4 * s + lk;
      a[s] = (s < ksteps && row < m && kk < k) ? A[row * lda + kk] : 0.0;
    }
    for (long g = wave_id; g < ngroups; g += nwaves) {
      dbl4 acc[kMulTiles];
#pragma unroll
      for (int t = 0; t < kMulTiles; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
      for (int s = 0; s < ksteps; ++s) {
        const int kk = 4 * s + lk;
#pragma unroll
        for (int t = 0; t < kMulTiles; ++t) {
          const long col = (g * kMulTiles + t) * 16 + li;
          const double b = (kk < k && col < ncol) ? B[kk * ldb + col] : 0.0;
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b, acc[t], 0, 0, 0);
        }
      }
#pragma unroll
      for (int t = 0; t < kMulTiles; ++t) {
        const long col = (g * kMulTiles + t) * 16 + li;
        if (col >= ncol) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int orow = rt * 16 + 4 * r + lk;
          if (orow < m) {
            double* cp = C + orow * ldc + col;
            const double v = alpha * acc[t][r];
            *cp = (beta == 0.0) ? v : v + beta * *cp;
          }
        }
      }
    }
  }
}

// Paired-column form for 16-byte aligned operands (B, C, ldb, ldc and ncol
// even): lane (li, lk) loads B[4s + lk][c0 + 2 li .. + 1] as one 16-byte load
// per 32-column quad q, feeding two 16 x 16 tiles, tile 2q with the even and
// tile 2q+1 with the odd columns, and stores the two tiles' results for a row
// as one 16-byte store.  A wave takes 128 columns (8 tiles) per group; KS =
// ceil(k/4) is a template parameter, so the A operand stays in registers and
// every k-step's loads can be in flight together.  Rows k..4KS-1 of the last
// step read row k-1 (A is zero there); only the tail group checks columns.
typedef double dbl2 __attribute__((ext_vector_type(2)));
constexpr int kMul2Cols = 128;

template <int KS>
__global__ __launch_bounds__(256) void k_mul_mfma2(const double* __restrict__ A, int m, int k, long lda,
                                                   const double* __restrict__ B, long ldb,
                                                   double* __restrict__ C, long ldc, long ncol,
                                                   double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int rtiles = (m + 15) / 16;
  const long nfull = ncol / kMul2Cols;
  const long ngroups = (ncol + kMul2Cols - 1) / kMul2Cols;
  const long wave_id = static_cast<long>(blockIdx.x) * (blockDim.x >> 6) + wave;
  const long nwaves = static_cast<long>(gridDim.x) * (blockDim.x >> 6);
  const double* brow[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) brow[s] = B + static_cast<long>(min(4 * s + lk, k - 1)) * ldb + 2 * li;

  for (int rt = 0; rt < rtiles; ++rt) {
    double a[KS];
    const int row = rt * 16 + li;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int kk = 4 * s + lk;
      a[s] = (row < m && kk < k) ? A[row * lda + kk] : 0.0;
    }
    for (long g = wave_id; g < ngroups; g += nwaves) {
      const long c0 = g * kMul2Cols;
      dbl4 acc[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
      const bool full = g < nfull;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        dbl2 b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long col = c0 + 32 * q + 2 * li;
          b[q] = (full || col < ncol) ? *reinterpret_cast<const dbl2*>(brow[s] + c0 + 32 * q) : dbl2{0.0, 0.0};
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc[2 * q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[q].x, acc[2 * q], 0, 0, 0);
          acc[2 * q + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[q].y, acc[2 * q + 1], 0, 0, 0);
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long col = c0 + 32 * q + 2 * li;
        if (!full && col >= ncol) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int orow = rt * 16 + 4 * r + lk;
          if (orow < m) {
            dbl2* cp = reinterpret_cast<dbl2*>(C + orow * ldc + col);
            dbl2 v = dbl2{alpha * acc[2 * q][r], alpha * acc[2 * q + 1][r]};
            if (beta != 0.0) v += beta * *cp;
            *cp = v;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// pointwise kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ long soa_ix(long e, long v, long nv) { return ((e >> 5) * nv + v) * 32 + (e & 31); }

// transformed advective flux, tflux.mako:6-25 with flux.mako:4-20:
// f[i][upt][v][e] = sum_k smats[i][upt][k][e] * c_k[vidx+v] * u[upt][v][e]
__global__ __launch_bounds__(256) void k_tflux(int ndims, int nupts, int nvars, int vidx, long ncol,
                                               const double* __restrict__ u, long ldu,
                                               const double* __restrict__ smats, long ldsm,
                                               double* __restrict__ f, long ldf,
                                               const double* __restrict__ cv, int vsize) {
  const long c = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int upt = blockIdx.y;
  if (c >= ncol) return;
  const long blk = c >> 5, lane = c & 31;
  const int v = static_cast<int>(blk % nvars);
  const long e = (blk / nvars) * 32 + lane;
  const double s = u[upt * ldu + c];
  double cvv[3];
  for (int d = 0; d < ndims; ++d) cvv[d] = cv[d * static_cast<long>(vsize) + vidx + v];
  for (int i = 0; i < ndims; ++i) {
    double acc = 0.0;
    for (int k2 = 0; k2 < ndims; ++k2)
      acc += smats[(static_cast<long>(i) * nupts + upt) * ldsm + soa_ix(e, k2, ndims)] * cvv[k2];
    f[(static_cast<long>(i) * nupts + upt) * ldf + c] = acc * s;
  }
}

// Paired forms of tflux and negdivconf for 16-byte aligned registers: a
// thread takes two neighbouring elements of one 32-element AoSoA block (one
// 16-byte load and store per row), a workgroup 16 consecutive blocks; the
// (element block, velocity) pair of the workgroup's first block comes from
// one division, the threads step from it.
constexpr int kPairBlocks = 16;

__device__ __forceinline__ void block_coords(long nvars, int j, long* eb, int* v) {
  const unsigned b0 = blockIdx.x * kPairBlocks;
  const unsigned e0 = b0 / static_cast<unsigned>(nvars);
  long e = e0;
  int vv = static_cast<int>(b0 - e0 * static_cast<unsigned>(nvars)) + j;
  while (vv >= nvars) {
    vv -= static_cast<int>(nvars);
    ++e;
  }
  *eb = e;
  *v = vv;
}

__global__ __launch_bounds__(256) void k_tflux2(int ndims, int nupts, int nvars, int vidx, long nblk,
                                                const double* __restrict__ u, long ldu,
                                                const double* __restrict__ smats, long ldsm,
                                                double* __restrict__ f, long ldf,
                                                const double* __restrict__ cv, int vsize) {
  const int j = threadIdx.x >> 4, lane = 2 * (threadIdx.x & 15);
  const long blk = static_cast<long>(blockIdx.x) * kPairBlocks + j;
  const int upt = blockIdx.y;
  if (blk >= nblk) return;
  long eb;
  int v;
  block_coords(nvars, j, &eb, &v);
  const long c = blk * 32 + lane;
  const dbl2 s = *reinterpret_cast<const dbl2*>(u + upt * ldu + c);
  double cvv[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < ndims; ++d) cvv[d] = cv[d * static_cast<long>(vsize) + vidx + v];
  for (int i = 0; i < ndims; ++i) {
    const double* sm = smats + (static_cast<long>(i) * nupts + upt) * ldsm + eb * ndims * 32 + lane;
    dbl2 acc = dbl2{0.0, 0.0};
    for (int k2 = 0; k2 < ndims; ++k2) acc += *reinterpret_cast<const dbl2*>(sm + 32 * k2) * cvv[k2];
    *reinterpret_cast<dbl2*>(f + (static_cast<long>(i) * nupts + upt) * ldf + c) = acc * s;
  }
}

__global__ __launch_bounds__(256) void k_negdivconf2(int nvars, int neles, long nblk, double* __restrict__ out,
                                                     long ld, const double* __restrict__ rcpdjac, long ldr) {
  const int j = threadIdx.x >> 4, lane = 2 * (threadIdx.x & 15);
  const long blk = static_cast<long>(blockIdx.x) * kPairBlocks + j;
  const int upt = blockIdx.y;
  if (blk >= nblk) return;
  long eb;
  int v;
  block_coords(nvars, j, &eb, &v);
  const long e = eb * 32 + lane;
  if (e >= neles) return;
  dbl2* p = reinterpret_cast<dbl2*>(out + upt * ld + blk * 32 + lane);
  const double r0 = rcpdjac[upt * ldr + e];
  const double r1 = (e + 1 < neles) ? rcpdjac[upt * ldr + e + 1] : 0.0;
  dbl2 x = *p;
  if (e + 1 < neles) {
    *p = dbl2{-r0 * x.x, -r1 * x.y};
  } else {
    out[upt * ld + blk * 32 + lane] = -r0 * x.x;
  }
}

// negdivconf.mako:5-14 with no source term: out = -rcpdjac * out.  rcpdjac
// is a plain (nupts, neles) matrix, so the padding elements of the last AoSoA
// block (e >= neles) are skipped.
__global__ __launch_bounds__(256) void k_negdivconf(int nvars, int neles, long ncol, double* __restrict__ out,
                                                    long ld, const double* __restrict__ rcpdjac,
                                                    long ldr) {
  const long c = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int upt = blockIdx.y;
  if (c >= ncol) return;
  const long e = ((c >> 5) / nvars) * 32 + (c & 31);
  if (e >= neles) return;
  double* p = out + upt * ld + c;
  *p = -rcpdjac[upt * ldr + e] * *p;
}

// exact upwind (Rusanov with a = |n.c|) common flux, rusanov.mako:47-49,
// written back as ul = |n| fn, ur = -|n| fn (intcflux.mako:21-23).
// Thread (i, velocity chunk); interface points are ordered so consecutive i
// touch consecutive elements of the left state.
// L and R address the same flux-point matrix on a one-type mesh (distinct
// elements), so they carry no restrict qualifier.
__global__ __launch_bounds__(256) void k_intcflux(int ndims, int nfpts, int nvars, int vidx, int vchunk,
                                                  double* L, const long* __restrict__ loff,
                                                  double* R, const long* __restrict__ roff,
                                                  const double* __restrict__ nl,
                                                  const double* __restrict__ magnl,
                                                  const double* __restrict__ cv, int vsize) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfpts) return;
  const int v0 = blockIdx.y * vchunk, v1 = min(v0 + vchunk, nvars);
  const long ol = loff[i], orr = roff[i];
  double n[3] = {0, 0, 0};
  for (int d = 0; d < ndims; ++d) n[d] = nl[d * static_cast<long>(nfpts) + i];
  const double mag = magnl[i];
  for (int v = v0; v < v1; ++v) {
    double nc = 0.0;
    for (int d = 0; d < ndims; ++d) nc += n[d] * cv[d * static_cast<long>(vsize) + vidx + v];
    const double ul = L[ol + 32L * v], ur = R[orr + 32L * v];
    const double fn = 0.5 * nc * (ul + ur) + 0.5 * fabs(nc) * (ul - ur);
    L[ol + 32L * v] = mag * fn;
    R[orr + 32L * v] = -mag * fn;
  }
}

// boundary types of bccflux.mako with bcs/*.mako
enum BcType {
  BC_DIFFUSE = 0, BC_SPECULAR = 1, BC_INLET = 2, BC_OUTLET = 3, BC_VANISH = 4, BC_INLET_NS = 5,
  BC_DIFFUSE_CYL = 6
};

// wall of dgfs-wall-diffuse-cyl.mako: per-point wall velocity and
// temperature st[4][nfpts] = (Ux, Uy, Uz, T) and the unnormalised wall
// Maxwellian exp(-|c - U|^2 / T) (its normalisation cancels in the flux
// balance ratio)
__device__ inline double cyl_wall_f(const double* cv, long vsize, int vv, const double U[3], double T) {
  const double dx = cv[vv] - U[0], dy = cv[vsize + vv] - U[1], dz = cv[2 * vsize + vv] - U[2];
  return exp(-(dx * dx + dy * dy + dz * dz) / T);
}

struct BcArgs {
  int bctype, ndims, nfpts, nvars, vidx, vchunk, vsize;
  double* L;
  const long* loff;
  const double* nl;
  const double* magnl;
  const double* f0;       // wall/inlet Maxwellian [vsize]; cylinder wall: st[4][nfpts]
  const double* bcvals;   // diffuse: [2][nfpts]; specular: stored ul [vsize][nfpts]
  const int* rmap;        // specular reflection map [vsize]
  const double* cv;
  double U[3];
};

// ghost state (bcs/dgfs-*.mako) then the upwind flux; ul = |n| fn
__global__ __launch_bounds__(256) void k_bccflux(BcArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nfpts) return;
  const int v0 = blockIdx.y * a.vchunk, v1 = min(v0 + a.vchunk, a.nvars);
  const long ol = a.loff[i];
  double n[3] = {0, 0, 0};
  for (int d = 0; d < a.ndims; ++d) n[d] = a.nl[d * static_cast<long>(a.nfpts) + i];
  const double mag = a.magnl[i];
  double ratio = 0.0, U[3] = {a.U[0], a.U[1], a.U[2]}, Tw = 1.0;
  if (a.bctype == BC_DIFFUSE || a.bctype == BC_DIFFUSE_CYL) ratio = -a.bcvals[i] / a.bcvals[a.nfpts + i];
  if (a.bctype == BC_DIFFUSE_CYL) {
    for (int d = 0; d < 3; ++d) U[d] = a.f0[d * static_cast<long>(a.nfpts) + i];
    Tw = a.f0[3L * a.nfpts + i];
  }
  for (int v = v0; v < v1; ++v) {
    const int vv = a.vidx + v;
    double nc = 0.0, fac = 0.0;
    for (int d = 0; d < a.ndims; ++d) {
      const double c = a.cv[d * static_cast<long>(a.vsize) + vv];
      nc += n[d] * c;
      fac += (c - U[d]) * n[d];
    }
    const double ul = a.L[ol + 32L * v];
    double ur = ul;
    switch (a.bctype) {
      case BC_DIFFUSE:   // dgfs-wall-diffuse.mako:13-17
        ur = (fac < 0) ? a.f0[vv] * ratio : ul;
        break;
      case BC_DIFFUSE_CYL:  // dgfs-wall-diffuse-cyl.mako:12-27
        ur = (fac < 0) ? cyl_wall_f(a.cv, a.vsize, vv, U, Tw) * ratio : ul;
        break;
      case BC_SPECULAR:  // dgfs-wall-specular.mako:12-15
        ur = a.bcvals[static_cast<long>(a.rmap[vv]) * a.nfpts + i];
        break;
      case BC_INLET:     // dgfs-inlet.mako:12-16
        ur = (fac < 0) ? a.f0[vv] : ul;
        break;
      case BC_INLET_NS:  // dgfs-inlet-normalshock.mako:12-16
        ur = (nc < 0) ? a.f0[vv] : ul;
        break;
      case BC_VANISH:    // dgfs-vanish.mako:9-13
        ur = (nc < 0) ? 0.0 : ul;
        break;
      default:           // dgfs-outlet.mako
        ur = ul;
        break;
    }
    const double fn = 0.5 * nc * (ul + ur) + 0.5 * fabs(nc) * (ul - ur);
    a.L[ol + 32L * v] = mag * fn;
  }
}

// diffuse-wall flux balance over the velocities of this batch
// (dgfs-wall-diffuse-update.mako:12-36): bc0 += sum_{fac>=0} fac ul cw,
// bc1 += sum_{fac<0} fac f0 cw; reset when vidx == 0.  With st set (the
// cylinder wall, dgfs-wall-diffuse-cyl-update.mako:12-44) U and T are
// per point and f0 is the unnormalised Maxwellian.  One block per wall
// point, a tree reduction over the velocity index.
__global__ __launch_bounds__(256) void k_bccupdate_diffuse(int ndims, int nfpts, int nvars, int vidx,
                                                           const double* __restrict__ L,
                                                           const long* __restrict__ loff,
                                                           const double* __restrict__ nl,
                                                           const double* __restrict__ f0,
                                                           double* __restrict__ bcvals, double Ux,
                                                           double Uy, double Uz, double cw,
                                                           const double* __restrict__ cv, int vsize,
                                                           const double* __restrict__ st) {
  __shared__ double r0[256], r1[256];
  const int i = blockIdx.x, tid = threadIdx.x;
  const long ol = loff[i];
  double n[3] = {0, 0, 0};
  for (int d = 0; d < ndims; ++d) n[d] = nl[d * static_cast<long>(nfpts) + i];
  double U[3] = {Ux, Uy, Uz}, Tw = 1.0;
  if (st) {
    for (int d = 0; d < 3; ++d) U[d] = st[d * static_cast<long>(nfpts) + i];
    Tw = st[3L * nfpts + i];
  }
  double s0 = 0.0, s1 = 0.0;
  for (int v = tid; v < nvars; v += 256) {
    const int vv = vidx + v;
    double fac = 0.0;
    for (int d = 0; d < ndims; ++d) fac += (cv[d * static_cast<long>(vsize) + vv] - U[d]) * n[d];
    if (fac >= 0) s0 += fac * L[ol + 32L * v] * cw;
    else s1 += fac * (st ? cyl_wall_f(cv, vsize, vv, U, Tw) : f0[vv]) * cw;
  }
  r0[tid] = s0;
  r1[tid] = s1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      r0[tid] += r0[tid + s];
      r1[tid] += r1[tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (vidx == 0) {
      bcvals[i] = r0[0];
      bcvals[nfpts + i] = r1[0];
    } else {
      bcvals[i] += r0[0];
      bcvals[nfpts + i] += r1[0];
    }
  }
}

// specular wall: keep a copy of the wall-side state for the reflected lookup
// (dgfs-wall-specular-update.mako:12-14)
__global__ __launch_bounds__(256) void k_bccupdate_specular(int nfpts, int nvars, int vidx,
                                                            const double* __restrict__ L,
                                                            const long* __restrict__ loff,
                                                            double* __restrict__ store) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfpts) return;
  const int v = blockIdx.y;
  store[static_cast<long>(vidx + v) * nfpts + i] = L[loff[i] + 32L * v];
}

// x0 = a0 x0 + sum_i a_i x_i (blasext axnpby.mako:5-36)
constexpr int kAxMax = 8;
struct AxArgs {
  double* x[kAxMax];
  double a[kAxMax];
  int nv;
  long n;
};

__global__ __launch_bounds__(256) void k_axnpby(AxArgs a) {
  const long stride = static_cast<long>(gridDim.x) * blockDim.x;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    double acc = (a.a[0] == 0.0) ? 0.0 : a.a[0] * a.x[0][i];
    for (int j = 1; j < a.nv; ++j) acc += a.a[j] * a.x[j][i];
    a.x[0][i] = acc;
  }
}

// velocity-slice copy between registers with nvsrc and nvdst variables
// (copy_dgfs.mako:5-26): dst[.., d0+v, ..] = src[.., s0+v, ..], v < count
__global__ __launch_bounds__(256) void k_copy_reg(int nblocks, int count, const double* __restrict__ src,
                                                  long lds, int nvsrc, int s0, double* __restrict__ dst,
                                                  long ldd, int nvdst, int d0) {
  const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int upt = blockIdx.y;
  const long total = static_cast<long>(nblocks) * count * 32;
  if (t >= total) return;
  const int lane = static_cast<int>(t & 31);
  const long q = t >> 5;
  const int v = static_cast<int>(q % count);
  const long eb = q / count;
  dst[upt * ldd + (eb * nvdst + d0 + v) * 32 + lane] = src[upt * lds + (eb * nvsrc + s0 + v) * 32 + lane];
}

// macroscopic moments per (upt, elem) over the whole velocity space
// (plugins/dgfsmomwriterstd.py:14-214): rho, U, T, q, P with
// T = (2/3) sum f |c-U|^2 cw / rho, q_i = sum f c'_i |c'|^2 cw, P_ij = 2 sum f c'_i c'_j cw,
// p = rho T.  out[upt][m][e] for m over the 15 outputs (3D layout).
__global__ __launch_bounds__(256) void k_moments(int nupts, int neles, int vsize, const double* __restrict__ f,
                                                 long ld, const double* __restrict__ cv, double cw,
                                                 double* __restrict__ out) {
  // one thread per (upt, e): two passes over v (bulk velocity, then peculiar moments)
  const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long>(nupts) * neles) return;
  const int upt = static_cast<int>(t / neles);
  const long e = t % neles;
  const double* fu = f + upt * ld;
  double rho = 0, ux = 0, uy = 0, uz = 0;
  for (int v = 0; v < vsize; ++v) {
    const double x = fu[soa_ix(e, v, vsize)];
    rho += x;
    ux += x * cv[v];
    uy += x * cv[vsize + v];
    uz += x * cv[2L * vsize + v];
  }
  rho *= cw;
  ux = ux * cw / rho;
  uy = uy * cw / rho;
  uz = uz * cw / rho;
  double T = 0, qx = 0, qy = 0, qz = 0, pxx = 0, pxy = 0, pxz = 0, pyy = 0, pyz = 0, pzz = 0;
  for (int v = 0; v < vsize; ++v) {
    const double x = fu[soa_ix(e, v, vsize)];
    const double cx = cv[v] - ux, cy = cv[vsize + v] - uy, cz = cv[2L * vsize + v] - uz;
    const double c2 = cx * cx + cy * cy + cz * cz;
    T += x * c2;
    qx += x * cx * c2;
    qy += x * cy * c2;
    qz += x * cz * c2;
    pxx += x * cx * cx; pxy += x * cx * cy; pxz += x * cx * cz;
    pyy += x * cy * cy; pyz += x * cy * cz; pzz += x * cz * cz;
  }
  T = T * cw * (2.0 / 3.0) / rho;
  const double vals[15] = {rho, ux, uy, uz, T, qx * cw, qy * cw, qz * cw, 2 * pxx * cw, 2 * pxy * cw,
                           2 * pxz * cw, 2 * pyy * cw, 2 * pyz * cw, 2 * pzz * cw, rho * T};
  for (int m = 0; m < 15; ++m) out[(static_cast<long>(upt) * 15 + m) * neles + e] = vals[m];
}

static bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

static dim3 grid2(long ncol, int rows) {
  return dim3(static_cast<unsigned>((ncol + 255) / 256), static_cast<unsigned>(rows));
}

}  // namespace dgfs

using namespace dgfs;

extern "C" int frfs_mul(const double* A, int m, int k, long lda, const double* B, long ldb, double* C,
                        long ldc, long ncol, double alpha, double beta, void* stream) {
  if (!A || !B || !C) return fail(DGFS_ERR_INVALID_VALUE, "frfs_mul: null operand");
  if (m < 1 || k < 1 || ncol < 0) return fail(DGFS_ERR_INVALID_VALUE, "frfs_mul: bad shape m=%d k=%d", m, k);
  if (k > kMulMaxK || m > 256) return fail(DGFS_ERR_UNSUPPORTED, "frfs_mul: k=%d > %d or m=%d > 256", k, kMulMaxK, m);
  if (ldb < ncol || ldc < ncol || lda < k)
    return fail(DGFS_ERR_INVALID_VALUE, "frfs_mul: leading dimension too small");
  if (ncol == 0) return DGFS_OK;
  auto st = static_cast<hipStream_t>(stream);
  const bool paired = (reinterpret_cast<uintptr_t>(B) % 16 == 0) && (reinterpret_cast<uintptr_t>(C) % 16 == 0) &&
                      ldb % 2 == 0 && ldc % 2 == 0 && ncol % 2 == 0;
  if (paired) {
    const long groups = (ncol + kMul2Cols - 1) / kMul2Cols;
    const unsigned blocks = static_cast<unsigned>(std::max<long>(1, std::min<long>((groups + 3) / 4, 8L * cu_count())));
    switch ((k + 3) / 4) {
#define DGFS_MUL2(KS) \
  case KS: hipLaunchKernelGGL(k_mul_mfma2<KS>, dim3(blocks), dim3(256), 0, st, A, m, k, lda, B, ldb, C, ldc, ncol, alpha, beta); break;
      DGFS_MUL2(1) DGFS_MUL2(2) DGFS_MUL2(3) DGFS_MUL2(4) DGFS_MUL2(5) DGFS_MUL2(6) DGFS_MUL2(7) DGFS_MUL2(8)
      DGFS_MUL2(9) DGFS_MUL2(10) DGFS_MUL2(11) DGFS_MUL2(12) DGFS_MUL2(13) DGFS_MUL2(14) DGFS_MUL2(15) DGFS_MUL2(16)
#undef DGFS_MUL2
      default: return fail(DGFS_ERR_UNSUPPORTED, "frfs_mul: k=%d", k);
    }
    DGFS_LAUNCH_CHECK("frfs_mul");
    return DGFS_OK;
  }
  const long ctiles = (ncol + 15) / 16;
  const long groups = (ctiles + kMulTiles - 1) / kMulTiles;
  const long blocks = std::min<long>((groups + 3) / 4, 8L * cu_count());
  hipLaunchKernelGGL(k_mul_mfma, dim3(static_cast<unsigned>(std::max<long>(blocks, 1))), dim3(256), 0, st, A, m, k,
                     lda, B, ldb, C, ldc, ncol, alpha, beta);
  DGFS_LAUNCH_CHECK("frfs_mul");
  return DGFS_OK;
}

extern "C" int dgfs_tflux(int ndims, int nupts, int neles, int nvars, int vidx, const double* u, long ldu,
                          const double* smats, long ldsm, double* f, long ldf, const double* cv, int vsize,
                          void* stream) {
  if (!u || !smats || !f || !cv) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_tflux: null pointer");
  if (ndims < 1 || ndims > 3 || vidx < 0 || vidx + nvars > vsize)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_tflux: bad ndims/velocity range");
  const long ncol = ((neles + 31L) / 32) * 32 * nvars;
  if (ldu < ncol || ldf < ncol) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_tflux: leading dimension < ncol");
  if (ncol == 0 || nupts == 0) return DGFS_OK;
  const long nblk = ncol / 32;
  if (aligned16(u) && aligned16(smats) && aligned16(f) && ldu % 2 == 0 && ldsm % 2 == 0 && ldf % 2 == 0 &&
      nblk < (1L << 31) / kPairBlocks) {
    hipLaunchKernelGGL(k_tflux2, dim3(static_cast<unsigned>((nblk + kPairBlocks - 1) / kPairBlocks), nupts),
                       dim3(256), 0, static_cast<hipStream_t>(stream), ndims, nupts, nvars, vidx, nblk, u, ldu,
                       smats, ldsm, f, ldf, cv, vsize);
  } else {
    hipLaunchKernelGGL(k_tflux, grid2(ncol, nupts), dim3(256), 0, static_cast<hipStream_t>(stream), ndims, nupts,
                       nvars, vidx, ncol, u, ldu, smats, ldsm, f, ldf, cv, vsize);
  }
  DGFS_LAUNCH_CHECK("dgfs_tflux");
  return DGFS_OK;
}

extern "C" int dgfs_negdivconf(int nupts, int neles, int nvars, double* out, long ld, const double* rcpdjac,
                               long ldr, void* stream) {
  if (!out || !rcpdjac) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_negdivconf: null pointer");
  const long ncol = ((neles + 31L) / 32) * 32 * nvars;
  if (ld < ncol || ldr < neles) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_negdivconf: leading dimension");
  if (ncol == 0 || nupts == 0) return DGFS_OK;
  const long nblk = ncol / 32;
  if (aligned16(out) && ld % 2 == 0 && nblk < (1L << 31) / kPairBlocks) {
    hipLaunchKernelGGL(k_negdivconf2, dim3(static_cast<unsigned>((nblk + kPairBlocks - 1) / kPairBlocks), nupts),
                       dim3(256), 0, static_cast<hipStream_t>(stream), nvars, neles, nblk, out, ld, rcpdjac, ldr);
  } else {
    hipLaunchKernelGGL(k_negdivconf, grid2(ncol, nupts), dim3(256), 0, static_cast<hipStream_t>(stream), nvars,
                       neles, ncol, out, ld, rcpdjac, ldr);
  }
  DGFS_LAUNCH_CHECK("dgfs_negdivconf");
  return DGFS_OK;
}

static int vchunk_for(int nvars) { return std::max(1, std::min(nvars, 64)); }

extern "C" int dgfs_intcflux(int ndims, int nfpts, int nvars, int vidx, double* L, const long* loff, double* R,
                             const long* roff, const double* nl, const double* magnl, const double* cv,
                             int vsize, void* stream) {
  if (!L || !loff || !R || !roff || !nl || !magnl || !cv)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_intcflux: null pointer");
  if (ndims < 1 || ndims > 3 || vidx < 0 || vidx + nvars > vsize)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_intcflux: bad ndims/velocity range");
  if (nfpts == 0 || nvars == 0) return DGFS_OK;
  const int vc = vchunk_for(nvars);
  dim3 grid((nfpts + 255) / 256, (nvars + vc - 1) / vc);
  hipLaunchKernelGGL(k_intcflux, grid, dim3(256), 0, static_cast<hipStream_t>(stream), ndims, nfpts, nvars,
                     vidx, vc, L, loff, R, roff, nl, magnl, cv, vsize);
  DGFS_LAUNCH_CHECK("dgfs_intcflux");
  return DGFS_OK;
}

extern "C" int dgfs_bccflux(int bctype, int ndims, int nfpts, int nvars, int vidx, double* L, const long* loff,
                            const double* nl, const double* magnl, const double* f0, const double* bcvals,
                            const int* rmap, double Ux, double Uy, double Uz, const double* cv, int vsize,
                            void* stream) {
  if (!L || !loff || !nl || !magnl || !cv) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccflux: null pointer");
  if (bctype < BC_DIFFUSE || bctype > BC_DIFFUSE_CYL) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccflux: bctype");
  if ((bctype == BC_DIFFUSE || bctype == BC_INLET || bctype == BC_INLET_NS || bctype == BC_DIFFUSE_CYL) && !f0)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccflux: this boundary needs f0");
  if ((bctype == BC_DIFFUSE || bctype == BC_SPECULAR || bctype == BC_DIFFUSE_CYL) && !bcvals)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccflux: this boundary needs bc_vals");
  if (bctype == BC_SPECULAR && !rmap) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccflux: specular needs rmap");
  if (ndims < 1 || ndims > 3 || vidx < 0 || vidx + nvars > vsize)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccflux: bad ndims/velocity range");
  if (nfpts == 0 || nvars == 0) return DGFS_OK;
  BcArgs a;
  a.bctype = bctype; a.ndims = ndims; a.nfpts = nfpts; a.nvars = nvars; a.vidx = vidx;
  a.vchunk = vchunk_for(nvars); a.vsize = vsize;
  a.L = L; a.loff = loff; a.nl = nl; a.magnl = magnl; a.f0 = f0; a.bcvals = bcvals; a.rmap = rmap; a.cv = cv;
  a.U[0] = Ux; a.U[1] = Uy; a.U[2] = Uz;
  dim3 grid((nfpts + 255) / 256, (nvars + a.vchunk - 1) / a.vchunk);
  hipLaunchKernelGGL(k_bccflux, grid, dim3(256), 0, static_cast<hipStream_t>(stream), a);
  DGFS_LAUNCH_CHECK("dgfs_bccflux");
  return DGFS_OK;
}

extern "C" int dgfs_bccupdate(int bctype, int ndims, int nfpts, int nvars, int vidx, const double* L,
                              const long* loff, const double* nl, const double* f0, double* bcvals,
                              double Ux, double Uy, double Uz, double cw, const double* cv, int vsize,
                              void* stream) {
  if (!L || !loff || !nl || !bcvals || !cv) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccupdate: null pointer");
  if (vidx < 0 || vidx + nvars > vsize) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccupdate: velocity range");
  if (nfpts == 0 || nvars == 0) return DGFS_OK;
  auto st = static_cast<hipStream_t>(stream);
  if (bctype == BC_DIFFUSE || bctype == BC_DIFFUSE_CYL) {
    if (!f0) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccupdate: diffuse wall needs f0");
    const bool cyl = bctype == BC_DIFFUSE_CYL;
    hipLaunchKernelGGL(k_bccupdate_diffuse, dim3(nfpts), dim3(256), 0, st, ndims, nfpts, nvars, vidx, L, loff,
                       nl, cyl ? nullptr : f0, bcvals, Ux, Uy, Uz, cw, cv, vsize, cyl ? f0 : nullptr);
  } else if (bctype == BC_SPECULAR) {
    hipLaunchKernelGGL(k_bccupdate_specular, dim3((nfpts + 255) / 256, nvars), dim3(256), 0, st, nfpts, nvars,
                       vidx, L, loff, bcvals);
  } else {
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_bccupdate: boundary type %d has no update", bctype);
  }
  DGFS_LAUNCH_CHECK("dgfs_bccupdate");
  return DGFS_OK;
}

extern "C" int frfs_axnpby(long n, int nv, double* const* x, const double* a, void* stream) {
  if (nv < 1 || nv > kAxMax || !x || !a) return fail(DGFS_ERR_INVALID_VALUE, "frfs_axnpby: 1..%d arrays", kAxMax);
  AxArgs args;
  args.nv = nv;
  args.n = n;
  for (int i = 0; i < kAxMax; ++i) {
    args.x[i] = (i < nv) ? x[i] : nullptr;
    args.a[i] = (i < nv) ? a[i] : 0.0;
    if (i < nv && !x[i]) return fail(DGFS_ERR_INVALID_VALUE, "frfs_axnpby: null array %d", i);
  }
  if (n <= 0) return DGFS_OK;
  const long blocks = std::min<long>((n + 255) / 256, 16L * cu_count());
  hipLaunchKernelGGL(k_axnpby, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     args);
  DGFS_LAUNCH_CHECK("frfs_axnpby");
  return DGFS_OK;
}

extern "C" int dgfs_copy_reg(int nupts, int neles, const double* src, long lds, int nvsrc, int s0, double* dst,
                             long ldd, int nvdst, int d0, int count, void* stream) {
  if (!src || !dst) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_copy_reg: null pointer");
  if (s0 < 0 || d0 < 0 || count < 0 || s0 + count > nvsrc || d0 + count > nvdst)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_copy_reg: velocity range outside the registers");
  const int nblocks = (neles + 31) / 32;
  if (lds < static_cast<long>(nblocks) * nvsrc * 32 || ldd < static_cast<long>(nblocks) * nvdst * 32)
    return fail(DGFS_ERR_INVALID_VALUE, "dgfs_copy_reg: leading dimension too small");
  const long total = static_cast<long>(nblocks) * count * 32;
  if (total == 0 || nupts == 0) return DGFS_OK;
  hipLaunchKernelGGL(k_copy_reg, dim3(static_cast<unsigned>((total + 255) / 256), nupts), dim3(256), 0,
                     static_cast<hipStream_t>(stream), nblocks, count, src, lds, nvsrc, s0, dst, ldd, nvdst, d0);
  DGFS_LAUNCH_CHECK("dgfs_copy_reg");
  return DGFS_OK;
}

extern "C" int dgfs_moments(int nupts, int neles, int vsize, const double* f, long ld, const double* cv, double cw,
                            double* out, void* stream) {
  if (!f || !cv || !out) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_moments: null pointer");
  const long nparr = ((neles + 31L) / 32) * 32;
  if (ld < nparr * vsize) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_moments: leading dimension");
  const long total = static_cast<long>(nupts) * neles;
  if (total == 0) return DGFS_OK;
  hipLaunchKernelGGL(k_moments, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), nupts, neles, vsize, f, ld, cv, cw, out);
  DGFS_LAUNCH_CHECK("dgfs_moments");
  return DGFS_OK;
}

// Residual of two full-velocity AoSoA registers (plugins/dgfsresidualstd.py:
// 41-49): out2 = (sum (curr - prev)^2, sum prev^2) over the valid entries
// (nupts rows, vsize variables, neles elements; AoSoA padding and row padding
// excluded).  Two deterministic stages: per-block partials, then one block;
// `work` holds 2 * kResidBlocks doubles.
namespace dgfs {
constexpr int kResidBlocks = 512;

__global__ __launch_bounds__(256) void k_resid_part(int nupts, int neles, int vsize, const double* __restrict__ prev,
                                                    const double* __restrict__ curr, long ld, double* part) {
  __shared__ double r0[256], r1[256];
  const long nb = (neles + 31) / 32;
  const long per_row = nb * vsize * 32;
  const long total = static_cast<long>(nupts) * per_row;
  double a = 0.0, b = 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const long row = i / per_row, j = i - row * per_row;
    const long eb = j / (static_cast<long>(vsize) * 32);
    if (eb * 32 + (j & 31) >= neles) continue;  // AoSoA padding of the last block
    const double p = prev[row * ld + j], c = curr[row * ld + j];
    a += (c - p) * (c - p);
    b += p * p;
  }
  r0[threadIdx.x] = a;
  r1[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      r0[threadIdx.x] += r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0[0];
    part[2 * blockIdx.x + 1] = r1[0];
  }
}

__global__ __launch_bounds__(64) void k_resid_final(const double* part, int nblk, double* out2) {
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int k = 0; k < nblk; ++k) s += part[2 * k + threadIdx.x];
    out2[threadIdx.x] = s;
  }
}
}  // namespace dgfs

extern "C" int dgfs_residual_norms(int nupts, int neles, int vsize, const double* prev, const double* curr, long ld,
                                   double* out2, double* work, void* stream) {
  using namespace dgfs;
  if (!prev || !curr || !out2 || !work) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_residual_norms: null pointer");
  const long nparr = ((neles + 31L) / 32) * 32;
  if (ld < nparr * vsize) return fail(DGFS_ERR_INVALID_VALUE, "dgfs_residual_norms: leading dimension");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_resid_part, dim3(kResidBlocks), dim3(256), 0, st, nupts, neles, vsize, prev, curr, ld, work);
  DGFS_LAUNCH_CHECK("dgfs_residual_norms");
  hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(64), 0, st, work, kResidBlocks, out2);
  DGFS_LAUNCH_CHECK("dgfs_residual_norms");
  return DGFS_OK;
}
