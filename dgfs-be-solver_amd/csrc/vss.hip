// Multi-species fast-spectral VSS collision operator Q_p = sum_q Q_pq(f_p, f_q)
// for gfx950.
//
// Dataflow of the reference's binary-mixture operator
// (frfs/solvers/dgfsbi/scattering.py:259-330, kernels/scattering/vss-gll.mako),
// per ordered species pair (p, q) and spatial point:
//
//   F = FFT(f_p)/n,  G = FFT(f_q)/n                              (prep kernel)
//   for every radial node x and full-sphere direction y, a = (pi/L) qz_x (s_y . l):
//     t3 = IFFT(exp(+i a m_q/(m_p+m_q)) F)                         (vss-gll.mako:148-151)
//     t4 = IFFT(exp(-i a m_p/(m_p+m_q)) G)                         (mako:153-156)
//     G^ += (qw_x sw / n) b_xy FFT(t3 t4)                          (mako:163-188, 224-250)
//   Q_p += prefac_pq (Re IFFT(G^) - f_p Re IFFT(c G))              (mako:252-300)
//
// The phase is separable, exp(i a) = ex(kx) ey(ky) ez(kz), so the reference's
// a table (Nrho*M*n) is never stored; b_xy (complex, Nrho*M*n per pair) and c
// (n per pair) are built on the device in FP64 from the Mpre-point design
// (vss-gll.mako:53-95).
//
// All three transforms of every (x, y) stay inside one workgroup: slab passes
// (kx-planes 2D-transformed in LDS) and pencil passes (x-pencils in registers)
// through two per-workgroup scratch cubes that the persistent grid reuses item
// after item, so they live in L2 / Infinity Cache.  The work items of all
// species pairs go into one launch.
//
// Antipodal sharing.  With a = (pi/L) qz_x (s_y . l) odd in s_y, the product
// of pair (q, p) at direction -s_y is the product of pair (p, q) at s_y, and
// the product of a pair (p, p) is even in s_y.  When the design holds every
// direction's antipode, one product per (x, y) therefore serves G^_pq with
// b_pq(x, y) and G^_qp with b_qp(x, -y), and a self pair runs over one
// direction of every antipodal pair with b(x, y) + b(x, -y): the transform
// count per point drops from 3 K S^2 to 3 K S (S + 1) / 2 ... 3 K S^2 / 2.
// The work is listed as jobs: a job names the two spectra, the phases and
// one or two (table, G^ output) pairs with their direction rows.
//
// Finalize: Re IFFT(Z) = IFFT(H(Z)) with H the Hermitian part,
// H(Z)(k) = (Z(k) + conj Z(-k))/2, so one complex transform of
// sum_q prefac_pq [H(G^_pq) + i H(c_pq G_q)] yields the gain (real part) and
// the loss field (imaginary part) of every pair of species p at once.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "cplx.hpp"
#include "spectral.hpp"

namespace dgfs {

constexpr int kVssMaxSpecies = 4;
constexpr int kVssMaxPairs = kVssMaxSpecies * kVssMaxSpecies;

// ---------------------------------------------------------------------------
// tables (FP64 on the device, stored in T)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lfreq(int k, int N) { return (k < N / 2) ? k : k - N; }

// b[x*M + y][r] = swpre qz_x^(gamma+2) sum_j B_yj exp(-i mpq (pi/L) qz_x (s'_j . l_r)),
// B_yj = (1 + s_y . s'_j)^eta (vss-gll.mako:59-94).
template <typename T>
__global__ __launch_bounds__(256) void k_vss_btab(cpx<T>* b, const double* __restrict__ Byj,
                                                  const double* __restrict__ szpre, const double* __restrict__ qz,
                                                  double L, double gamma, double mpq, double swpre, int N,
                                                  int M, int Mpre) {
  const int n = N * N * N, r = blockIdx.x * blockDim.x + threadIdx.x, xy = blockIdx.y;
  if (r >= n) return;
  const int x = xy / M, y = xy - x * M;
  const double lx = lfreq(r / (N * N), N), ly = lfreq((r / N) % N, N), lz = lfreq(r % N, N);
  const double fac = 3.14159265358979323846 / L * qz[x];
  const double fb = swpre * pow(qz[x], gamma + 2.0);
  double re = 0.0, im = 0.0;
  for (int j = 0; j < Mpre; ++j) {
    const double th = mpq * fac * (szpre[3 * j] * lx + szpre[3 * j + 1] * ly + szpre[3 * j + 2] * lz);
    double s, c;
    sincos(th, &s, &c);
    const double B = Byj[y * Mpre + j];
    re += B * fb * c;
    im -= B * fb * s;
  }
  b[static_cast<size_t>(xy) * n + r] = cpx<T>{static_cast<T>(re), static_cast<T>(im)};
}

// c[r] = sum_x qw_x sw swpre qz_x^(gamma+2) sum_j Bsum_j exp(-i (pi/L) qz_x (s'_j . l_r)),
// Bsum_j = sum_y B_yj (vss-gll.mako:86-92, accumulated over x, y).
template <typename T>
__global__ __launch_bounds__(256) void k_vss_ctab(cpx<T>* c, const double* __restrict__ Bsum,
                                                  const double* __restrict__ szpre, const double* __restrict__ qz,
                                                  const double* __restrict__ qw, double L, double gamma, double sw,
                                                  double swpre, int N, int Nrho, int Mpre) {
  const int n = N * N * N, r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double lx = lfreq(r / (N * N), N), ly = lfreq((r / N) % N, N), lz = lfreq(r % N, N);
  double re = 0.0, im = 0.0;
  for (int x = 0; x < Nrho; ++x) {
    const double fac = 3.14159265358979323846 / L * qz[x];
    const double fc = qw[x] * sw * swpre * pow(qz[x], gamma + 2.0);
    for (int j = 0; j < Mpre; ++j) {
      const double th = fac * (szpre[3 * j] * lx + szpre[3 * j + 1] * ly + szpre[3 * j + 2] * lz);
      double s, cc;
      sincos(th, &s, &cc);
      re += Bsum[j] * fc * cc;
      im -= Bsum[j] * fc * s;
    }
  }
  c[r] = cpx<T>{static_cast<T>(re), static_cast<T>(im)};
}

// ---------------------------------------------------------------------------
// gain: one work item = (pair, point, radial group)
// ---------------------------------------------------------------------------
// one job: products of spectra sp (phase mu1) and sq (phase -mu2) over
// ndir directions dirs[dir_off + i]; G^_out1 += w tab1[x][rows1[.]] FFT(.),
// and with out2 >= 0 also G^_out2 += w tab2[x][rows2[.]] FFT(.)
template <typename T>
struct VssJob {
  const cpx<T>* tab1;  // [Nrho][tabM1][n]
  const cpx<T>* tab2;  // [Nrho][tabM2][n] or null
  double mu1, mu2;     // m_q/(m_p+m_q), m_p/(m_p+m_q)
  int sp, sq, out1, out2, tabM1, tabM2, ndir, dir_off;
};

template <typename T>
struct VssGainArgs {
  const cpx<T>* F;   // [nspcs][npts][n] spectra, scaled by 1/n
  cpx<T>* Ghat;      // [npairs][npts][ngroups][n], pair index p * nspcs + q
  cpx<T>* Z1;        // scratch [nslots][n]
  cpx<T>* Z2;        // scratch [nslots][n]
  cpx<T>* Z3;        // scratch [nslots][n]: sum over directions (isotropic tables)
  const double* alpha;  // [Nrho] (pi/L) qz_x
  const double* wgt;    // [Nrho] qw_x sw / n
  const double* sz;     // [M][3]
  const int* dirs;      // direction, row-1 and row-2 lists of every job
  const int* rows1;
  const int* rows2;
  VssJob<T> job[kVssMaxPairs];
  int nitems, npts, ngroups, ppg, Nrho;
  int iso;              // eta = 0 everywhere: sum the products over directions, one forward transform per x
};

template <typename T, int N, bool ISO>
__global__ __launch_bounds__(Shape<N>::NT) void k_vss_gain(VssGainArgs<T> a) {
  constexpr int NT = Shape<N>::NT, PB = Shape<N>::PB, NN = N * N, n = N * N * N, h = N / 2;
  __shared__ cpx<T> lds[PB * N * (N + 1)];
  __shared__ cpx<T> e1[3][N], e2[3][N];
  const int tid = threadIdx.x;
  cpx<T>* Z1 = a.Z1 + static_cast<size_t>(blockIdx.x) * n;
  cpx<T>* Z2 = a.Z2 + static_cast<size_t>(blockIdx.x) * n;
  cpx<T>* Z3 = a.Z3 + static_cast<size_t>(blockIdx.x) * n;
  const int per_job = a.npts * a.ngroups;

  for (int item = blockIdx.x; item < a.nitems; item += gridDim.x) {
    const int j = item / per_job, rem = item - j * per_job;
    const int pt = rem / a.ngroups, g = rem - pt * a.ngroups;
    const int x0 = g * a.ppg, x1 = min(x0 + a.ppg, a.Nrho);
    const VssJob<T>& jb = a.job[j];
    const cpx<T>* F1 = a.F + (static_cast<size_t>(jb.sp) * a.npts + pt) * n;
    const cpx<T>* F2 = a.F + (static_cast<size_t>(jb.sq) * a.npts + pt) * n;
    const double mu1 = jb.mu1, mu2 = jb.mu2;
    cpx<T>* G = a.Ghat + ((static_cast<size_t>(jb.out1) * a.npts + pt) * a.ngroups + g) * n;
    cpx<T>* G2 = (jb.out2 >= 0) ? a.Ghat + ((static_cast<size_t>(jb.out2) * a.npts + pt) * a.ngroups + g) * n
                                : nullptr;
    for (int i = tid; i < n; i += NT) G[i] = cpx<T>{T(0), T(0)};
    if (G2)
      for (int i = tid; i < n; i += NT) G2[i] = cpx<T>{T(0), T(0)};
    for (int x = x0; x < x1; ++x) {
      const double alpha = a.alpha[x];
      const T w = static_cast<T>(a.wgt[x]);
      for (int iy = 0; iy < jb.ndir; ++iy) {
        const int y = a.dirs[jb.dir_off + iy];
        __syncthreads();
        if (tid < 3 * N) {
          const int d = tid / N, k = tid - d * N;
          const double th = alpha * a.sz[y * 3 + d] * lfreq(k, N);
          double s, c;
          sincos(mu1 * th, &s, &c);
          e1[d][k] = cpx<T>{static_cast<T>(c), static_cast<T>(s)};
          sincos(mu2 * th, &s, &c);
          e2[d][k] = cpx<T>{static_cast<T>(c), static_cast<T>(-s)};
        }
        __syncthreads();
        // t3 = IFFT(E1 F) into Z1
        for (int kx0 = 0; kx0 < N; kx0 += PB) {
          const int cnt = min(PB, N - kx0);
          slab_batch<T, N, NT, +1>(
              lds, cnt,
              [&](int j, int ky, int kz) -> cpx<T> {
                const int kx = kx0 + j;
                return F1[kx * NN + ky * N + kz] * (e1[0][kx] * e1[1][ky] * e1[2][kz]);
              },
              [&](int j, int yy, int z, cpx<T> v) { Z2[(kx0 + j) * NN + yy * N + z] = v; });
        }
        pencil_pass<T, N, NT, +1>([&](int kx, int idx) { return Z2[kx * NN + idx]; },
                                  [&](int idx, cpx<T>* v) {
#pragma unroll
                                    for (int xx = 0; xx < N; ++xx) Z1[xx * NN + idx] = v[xx];
                                  });
        // t4 = IFFT(E2 G); S = t3 t4 into Z1
        for (int kx0 = 0; kx0 < N; kx0 += PB) {
          const int cnt = min(PB, N - kx0);
          slab_batch<T, N, NT, +1>(
              lds, cnt,
              [&](int j, int ky, int kz) -> cpx<T> {
                const int kx = kx0 + j;
                return F2[kx * NN + ky * N + kz] * (e2[0][kx] * e2[1][ky] * e2[2][kz]);
              },
              [&](int j, int yy, int z, cpx<T> v) { Z2[(kx0 + j) * NN + yy * N + z] = v; });
        }
        if constexpr (ISO) {
          // S_x += t3 t4 (thread-private pencils: the same idx mapping in every pass)
          if (iy == 0) {
            pencil_pass<T, N, NT, +1>([&](int kx, int idx) { return Z2[kx * NN + idx]; },
                                      [&](int idx, cpx<T>* v) {
#pragma unroll
                                        for (int xx = 0; xx < N; ++xx) Z3[xx * NN + idx] = Z1[xx * NN + idx] * v[xx];
                                      });
          } else {
            pencil_pass<T, N, NT, +1>([&](int kx, int idx) { return Z2[kx * NN + idx]; },
                                      [&](int idx, cpx<T>* v) {
#pragma unroll
                                        for (int xx = 0; xx < N; ++xx) {
                                          const cpx<T> t = Z1[xx * NN + idx] * v[xx];
                                          Z3[xx * NN + idx] = Z3[xx * NN + idx] + t;
                                        }
                                      });
          }
          continue;
        }
        pencil_pass<T, N, NT, +1>([&](int kx, int idx) { return Z2[kx * NN + idx]; },
                                  [&](int idx, cpx<T>* v) {
#pragma unroll
                                    for (int xx = 0; xx < N; ++xx) Z1[xx * NN + idx] = Z1[xx * NN + idx] * v[xx];
                                  });
        // G^ += w b_xy FFT(S)
        pencil_pass<T, N, NT, -1>([&](int xx, int idx) { return Z1[xx * NN + idx]; },
                                  [&](int idx, cpx<T>* v) {
#pragma unroll
                                    for (int k = 0; k < N; ++k) Z2[k * NN + idx] = v[k];
                                  });
        const cpx<T>* bxy = jb.tab1 + (static_cast<size_t>(x) * jb.tabM1 + a.rows1[jb.dir_off + iy]) * n;
        if (G2) {
          const cpx<T>* bxy2 = jb.tab2 + (static_cast<size_t>(x) * jb.tabM2 + a.rows2[jb.dir_off + iy]) * n;
          for (int kx0 = 0; kx0 < N; kx0 += PB) {
            const int cnt = min(PB, N - kx0);
            slab_batch<T, N, NT, -1>(
                lds, cnt, [&](int jj, int r, int c) { return Z2[(kx0 + jj) * NN + r * N + c]; },
                [&](int jj, int r, int c, cpx<T> v) {
                  const int o = (kx0 + jj) * NN + r * N + c;
                  G[o] = G[o] + (w * bxy[o]) * v;
                  G2[o] = G2[o] + (w * bxy2[o]) * v;
                });
          }
        } else {
          for (int kx0 = 0; kx0 < N; kx0 += PB) {
            const int cnt = min(PB, N - kx0);
            slab_batch<T, N, NT, -1>(
                lds, cnt, [&](int jj, int r, int c) { return Z2[(kx0 + jj) * NN + r * N + c]; },
                [&](int jj, int r, int c, cpx<T> v) {
                  const int o = (kx0 + jj) * NN + r * N + c;
                  G[o] = G[o] + (w * bxy[o]) * v;
                });
          }
        }
      }
      if constexpr (ISO) {
        // G^ += w b(x) FFT(S_x), the tables read at the job's first direction row
        pencil_pass<T, N, NT, -1>([&](int xx, int idx) { return Z3[xx * NN + idx]; },
                                  [&](int idx, cpx<T>* v) {
#pragma unroll
                                    for (int k = 0; k < N; ++k) Z2[k * NN + idx] = v[k];
                                  });
        const cpx<T>* bx = jb.tab1 + (static_cast<size_t>(x) * jb.tabM1 + a.rows1[jb.dir_off]) * n;
        const cpx<T>* bx2 = G2 ? jb.tab2 + (static_cast<size_t>(x) * jb.tabM2 + a.rows2[jb.dir_off]) * n : nullptr;
        for (int kx0 = 0; kx0 < N; kx0 += PB) {
          const int cnt = min(PB, N - kx0);
          slab_batch<T, N, NT, -1>(
              lds, cnt, [&](int jj, int r, int c) { return Z2[(kx0 + jj) * NN + r * N + c]; },
              [&](int jj, int r, int c, cpx<T> v) {
                const int o = (kx0 + jj) * NN + r * N + c;
                G[o] = G[o] + (w * bx[o]) * v;
                if (bx2) G2[o] = G2[o] + (w * bx2[o]) * v;
              });
        }
      }
    }
  }
  (void)h;
}

// ---------------------------------------------------------------------------
// Register-resident gain, one pencil per thread (FP32, N = 16, 24: the N^3
// field fits in LDS).  A workgroup of N*N threads owns one work item; thread
// t = g N + l holds one pencil.  Per direction product:
//   t3: y-pencil (kx = g, kz = l) of F_p times the phase, y-IFFT; the kx
//       plane through LDS ([y][kz], rows padded to N+1) gives the z-pencil
//       (kx = g, y = l), z-IFFT; the field through LDS once ([kx][z][y])
//       gives the x-pencil (y = l, z = g), x-IFFT -- kept in registers;
//   t4: the same for F_q with the conjugate-side phase;
//   S = t3 t4 in registers; the reverse sequence (x, then z through the
//   field, then y through the plane) forward transforms S to y-pencils
//   (kx = g, kz = l) and adds w b FFT(S) to this thread's pencil of G^ (and
//   of the second output of a shared cross-pair job).
// The pencils that touch global memory (F, b, G^) run along ky, so at a
// fixed ky the lanes of a wave address consecutive values.  Same jobs,
// tables and arithmetic as k_vss_gain; no global scratch.
// ---------------------------------------------------------------------------
template <int N>
struct VssRR1Shape {
  static constexpr int NT = N * N;
  static constexpr int PP = N + 1;          // padded row pitch of the kx planes
  static constexpr int field = N * N * PP;  // complex; holds the N^3 exchange too
  // kx stride of the field exchange: N (N + 1) complex, so that the rows of
  // consecutive kx that one wave touches land on different LDS banks (a
  // stride of N^2, a multiple of 32 complex at N = 16 and 24, puts them on
  // the same banks); the exchange then fills the plane buffer exactly
  static constexpr int KS = N * PP;
  static_assert(field >= (N - 1) * KS + N * N, "exchange must fit the plane buffer");
  // values of t3 parked in LDS while t4 is transformed (frees registers where
  // the 576-thread workgroup's 168-VGPR limit would spill; the rest of the
  // 160 KiB of LDS holds them)
  static constexpr int park = (N == 24) ? 10 : 0;
};

// a pencil of stride N complex values in global memory addressed from two
// bases with immediate offsets (the offsets of one base stay under 4 KiB);
// the second base's distance is opaque to the compiler, so it does not fold
// the two bases back into one address per element
template <typename T, int N>
struct StridedPencil {
  static constexpr int KB = (4096 / static_cast<int>(N * sizeof(cpx<T>))) < N
                                ? (4096 / static_cast<int>(N * sizeof(cpx<T>))) : N;
  cpx<T>* p0;
  cpx<T>* p1;
  __device__ __forceinline__ explicit StridedPencil(cpx<T>* base) : p0(base) {
    int second = KB * N;
    asm volatile("" : "+v"(second));
    p1 = base + second;
  }
  __device__ __forceinline__ cpx<T>& operator[](int k) const { return (k < KB) ? p0[k * N] : p1[(k - KB) * N]; }
};

// inverse transform of a phased y-pencil (kx = g, kz = l) already in v:
// y-IFFT, plane exchange, z-IFFT, field exchange, x-IFFT; returns the
// x-pencil (x, y = l, z = g)
template <typename T, int N>
__device__ __forceinline__ void vss_rr1_inverse_tail(cpx<T> (&v)[N], cpx<T>* X, cpx<T>* Xp, int g, int l);

template <typename T, int N>
__device__ __forceinline__ void vss_rr1_inverse(cpx<T> (&v)[N], const cpx<T>* __restrict__ Fp, const cpx<T> (*e)[N],
                                                cpx<T>* X, cpx<T>* Xp, int g, int l) {
  constexpr int NN = N * N;
  const StridedPencil<T, N> src(const_cast<cpx<T>*>(Fp) + g * NN + l);
#pragma unroll
  for (int ky = 0; ky < N; ++ky) v[ky] = src[ky];
  const cpx<T> exz = e[0][g] * e[2][l];
#pragma unroll
  for (int ky = 0; ky < N; ++ky) v[ky] = v[ky] * (exz * e[1][ky]);
  vss_rr1_inverse_tail<T, N>(v, X, Xp, g, l);
}

template <typename T, int N>
__device__ __forceinline__ void vss_rr1_inverse_tail(cpx<T> (&v)[N], cpx<T>* X, cpx<T>* Xp, int g, int l) {
  constexpr int PP = VssRR1Shape<N>::PP, KS = VssRR1Shape<N>::KS;
  FFT<T, N, +1, true>::run(v);  // y: (kx = g, y, kz = l)
#pragma unroll
  for (int y = 0; y < N; ++y) Xp[y * PP + l] = v[y];
  __syncthreads();
#pragma unroll
  for (int kz = 0; kz < N; ++kz) v[kz] = Xp[l * PP + kz];
  FFT<T, N, +1, true>::run(v);  // z: (kx = g, y = l, z)
  __syncthreads();
#pragma unroll
  for (int z = 0; z < N; ++z) X[g * KS + z * N + l] = v[z];
  __syncthreads();
#pragma unroll
  for (int kx = 0; kx < N; ++kx) v[kx] = X[kx * KS + g * N + l];
  FFT<T, N, +1, true>::run(v);  // x: (x, y = l, z = g)
}

template <typename T, int N>
__global__ __launch_bounds__(VssRR1Shape<N>::NT) void k_vss_gain_rr1(VssGainArgs<T> a) {
  constexpr int NT = VssRR1Shape<N>::NT, PP = VssRR1Shape<N>::PP, NN = N * N, n = NN * N;
  constexpr int KS = VssRR1Shape<N>::KS;
  constexpr int KP = VssRR1Shape<N>::park;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X = reinterpret_cast<cpx<T>*>(smem_raw);
  __shared__ cpx<T> e1[3][N], e2[3][N];
  __shared__ cpx<T> upark[KP > 0 ? KP : 1][NT];
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  cpx<T>* Xp = X + g * N * PP;  // kx plane g
  const int per_job = a.npts * a.ngroups;
  cpx<T> u[N], v[N];
  for (int item = xcd_slot(); item < a.nitems; item += gridDim.x) {
    const int j = item / per_job, rem = item - j * per_job;
    const int pt = rem / a.ngroups, grp = rem - pt * a.ngroups;
    const int x0 = grp * a.ppg, x1 = min(x0 + a.ppg, a.Nrho);
    const VssJob<T>& jb = a.job[j];
    const cpx<T>* F1 = a.F + (static_cast<size_t>(jb.sp) * a.npts + pt) * n;
    const cpx<T>* F2 = a.F + (static_cast<size_t>(jb.sq) * a.npts + pt) * n;
    const double mu1 = jb.mu1, mu2 = jb.mu2;
    const size_t o = static_cast<size_t>(g) * NN + l;  // this thread's G^ pencil (kx = g, kz = l), stride N
    const bool two = jb.out2 >= 0;
    const StridedPencil<T, N> G(a.Ghat + ((static_cast<size_t>(jb.out1) * a.npts + pt) * a.ngroups + grp) * n + o);
    const StridedPencil<T, N> G2(a.Ghat + ((static_cast<size_t>(two ? jb.out2 : jb.out1) * a.npts + pt) *
                                               a.ngroups + grp) * n + o);
#pragma unroll
    for (int k = 0; k < N; ++k) G[k] = cpx<T>{T(0), T(0)};
    if (two) {
#pragma unroll
      for (int k = 0; k < N; ++k) G2[k] = cpx<T>{T(0), T(0)};
    }
    for (int x = x0; x < x1; ++x) {
      const double alpha = a.alpha[x];
      const T w = static_cast<T>(a.wgt[x]);
      for (int iy = 0; iy < jb.ndir; ++iy) {
        const int y = a.dirs[jb.dir_off + iy];
        __syncthreads();  // readers of e1, e2 and the LDS field of the previous product are done
        if (tid < 3 * N) {
          const int d = tid / N, k = tid - d * N;
          const double th = alpha * a.sz[y * 3 + d] * lfreq(k, N);
          double sn, cs;
          sincos(mu1 * th, &sn, &cs);
          e1[d][k] = cpx<T>{static_cast<T>(cs), static_cast<T>(sn)};
          sincos(mu2 * th, &sn, &cs);
          e2[d][k] = cpx<T>{static_cast<T>(cs), static_cast<T>(-sn)};
        }
        __syncthreads();
        vss_rr1_inverse<T, N>(u, F1, e1, X, Xp, g, l);  // t3 (x-pencil at y = l, z = g)
#pragma unroll
        for (int k = 0; k < KP; ++k) upark[k][tid] = u[N - KP + k];  // thread-private slots
        __syncthreads();  // every x-pencil of t3 is read before the planes are rewritten
        vss_rr1_inverse<T, N>(v, F2, e2, X, Xp, g, l);  // t4
#pragma unroll
        for (int k = 0; k < KP; ++k) u[N - KP + k] = upark[k][tid];
#pragma unroll
        for (int xx = 0; xx < N; ++xx) u[xx] = u[xx] * v[xx];
        FFT<T, N, -1, true>::run(u);  // x: (kx, y = l, z = g)
        __syncthreads();
#pragma unroll
        for (int kx = 0; kx < N; ++kx) X[kx * KS + g * N + l] = u[kx];
        __syncthreads();
#pragma unroll
        for (int z = 0; z < N; ++z) v[z] = X[g * KS + z * N + l];
        FFT<T, N, -1, true>::run(v);  // z: (kx = g, y = l, kz)
        __syncthreads();
#pragma unroll
        for (int kz = 0; kz < N; ++kz) Xp[kz * PP + l] = v[kz];
        __syncthreads();
#pragma unroll
        for (int yy = 0; yy < N; ++yy) v[yy] = Xp[l * PP + yy];
        FFT<T, N, -1, true>::run(v);  // y: (kx = g, ky, kz = l)
        {
          const StridedPencil<T, N> b1(const_cast<cpx<T>*>(jb.tab1) +
                                       (static_cast<size_t>(x) * jb.tabM1 + a.rows1[jb.dir_off + iy]) * n + o);
          pencil_accumulate<T, N, (N <= 16 ? 8 : 4)>(G, b1, w, v, false);
        }
        if (two) {
          const StridedPencil<T, N> b2(const_cast<cpx<T>*>(jb.tab2) +
                                       (static_cast<size_t>(x) * jb.tabM2 + a.rows2[jb.dir_off + iy]) * n + o);
          pencil_accumulate<T, N, (N <= 16 ? 8 : 4)>(G2, b2, w, v, false);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Register-resident gain for isotropic scattering (eta = 0 for every pair,
// the VHS mixture): B_yj = 1 makes b_pq(x, y) the same for every direction,
// so G^_pq = sum_x w_x b_pq(x) FFT(sum_y t3 t4) -- the direction products
// are summed in space and forward transformed once per radial node, as the
// single-species operator sums before its forward transform.  Per product:
// t3 (x-pencil) goes to this workgroup's scratch cube (thread-private),
// t4 is formed, and acc += t3 t4 in registers; per radial
// node the reverse sequence transforms acc to y-pencils (kx, kz) and adds
// w b(x) FFT(acc) to G^ (both outputs of a shared cross-pair job: the sum
// over all directions of the (p, q) products is also the (q, p) sum).  The
// tables are the general ones read at the job's first direction row.
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(VssRR1Shape<N>::NT) void k_vss_gain_rr1_iso(VssGainArgs<T> a) {
  constexpr int NT = VssRR1Shape<N>::NT, PP = VssRR1Shape<N>::PP, NN = N * N, n = NN * N;
  constexpr int KS = VssRR1Shape<N>::KS;
  constexpr int KP = VssRR1Shape<N>::park;  // accumulator values kept in the spare LDS
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X = reinterpret_cast<cpx<T>*>(smem_raw);
  __shared__ cpx<T> e1[3][N], e2[3][N];
  __shared__ cpx<T> apark[KP > 0 ? KP : 1][NT];
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  cpx<T>* Xp = X + g * N * PP;  // kx plane g
  // t3 parking: thread-contiguous pencils (one base register, immediate offsets)
  cpx<T>* U = a.Z1 + static_cast<size_t>(blockIdx.x) * n + static_cast<size_t>(tid) * N;
  // items run point-major: the jobs of one point sit on consecutive slots,
  // which xcd_slot() keeps on one XCD, so the point's spectra (read for
  // every product) are shared in that L2
  const int per_pt = (a.nitems / (a.npts * a.ngroups)) * a.ngroups;
  cpx<T> acc[N], v[N];
  for (int item = xcd_slot(); item < a.nitems; item += gridDim.x) {
    const int pt = item / per_pt, rem = item - pt * per_pt;
    const int j = rem / a.ngroups, grp = rem - j * a.ngroups;
    const int x0 = grp * a.ppg, x1 = min(x0 + a.ppg, a.Nrho);
    const VssJob<T>& jb = a.job[j];
    const cpx<T>* F1 = a.F + (static_cast<size_t>(jb.sp) * a.npts + pt) * n;
    const cpx<T>* F2 = a.F + (static_cast<size_t>(jb.sq) * a.npts + pt) * n;
    const double mu1 = jb.mu1, mu2 = jb.mu2;
    const size_t o = static_cast<size_t>(g) * NN + l;  // this thread's G^ pencil (kx = g, kz = l), stride N
    const bool two = jb.out2 >= 0;
    const StridedPencil<T, N> G(a.Ghat + ((static_cast<size_t>(jb.out1) * a.npts + pt) * a.ngroups + grp) * n + o);
    const StridedPencil<T, N> G2(a.Ghat + ((static_cast<size_t>(two ? jb.out2 : jb.out1) * a.npts + pt) *
                                               a.ngroups + grp) * n + o);
    for (int x = x0; x < x1; ++x) {
      const double alpha = a.alpha[x];
#pragma unroll
      for (int k = 0; k < N - KP; ++k) acc[k] = cpx<T>{T(0), T(0)};
#pragma unroll
      for (int k = 0; k < KP; ++k) apark[k][tid] = cpx<T>{T(0), T(0)};  // thread-private slots
      for (int iy = 0; iy < jb.ndir; ++iy) {
        const int y = a.dirs[jb.dir_off + iy];
        __syncthreads();  // readers of e1, e2 and the LDS field of the previous product are done
        if (tid < 3 * N) {
          const int d = tid / N, k = tid - d * N;
          const double th = alpha * a.sz[y * 3 + d] * lfreq(k, N);
          double sn, cs;
          sincos(mu1 * th, &sn, &cs);
          e1[d][k] = cpx<T>{static_cast<T>(cs), static_cast<T>(sn)};
          sincos(mu2 * th, &sn, &cs);
          e2[d][k] = cpx<T>{static_cast<T>(cs), static_cast<T>(-sn)};
        }
        __syncthreads();
        vss_rr1_inverse<T, N>(v, F1, e1, X, Xp, g, l);  // t3 (x-pencil at y = l, z = g)
#pragma unroll
        for (int xx = 0; xx < N; ++xx) U[xx] = v[xx];
        __syncthreads();  // every x-pencil of t3 is read before the planes are rewritten
        vss_rr1_inverse<T, N>(v, F2, e2, X, Xp, g, l);  // t4
#pragma unroll
        for (int xx = 0; xx < N - KP; ++xx) acc[xx] = acc[xx] + U[xx] * v[xx];
#pragma unroll
        for (int k = 0; k < KP; ++k) apark[k][tid] = apark[k][tid] + U[N - KP + k] * v[N - KP + k];
      }
#pragma unroll
      for (int k = 0; k < KP; ++k) acc[N - KP + k] = apark[k][tid];
      // G^ += w b(x) FFT(sum_y t3 t4)
      FFT<T, N, -1, true>::run(acc);  // x: (kx, y = l, z = g)
      __syncthreads();
#pragma unroll
      for (int kx = 0; kx < N; ++kx) X[kx * KS + g * N + l] = acc[kx];
      __syncthreads();
#pragma unroll
      for (int z = 0; z < N; ++z) v[z] = X[g * KS + z * N + l];
      FFT<T, N, -1, true>::run(v);  // z: (kx = g, y = l, kz)
      __syncthreads();
#pragma unroll
      for (int kz = 0; kz < N; ++kz) Xp[kz * PP + l] = v[kz];
      __syncthreads();
#pragma unroll
      for (int yy = 0; yy < N; ++yy) v[yy] = Xp[l * PP + yy];
      FFT<T, N, -1, true>::run(v);  // y: (kx = g, ky, kz = l)
      const T w = static_cast<T>(a.wgt[x]);
      {
        const StridedPencil<T, N> b1(const_cast<cpx<T>*>(jb.tab1) +
                                     (static_cast<size_t>(x) * jb.tabM1 + a.rows1[jb.dir_off]) * n + o);
        pencil_accumulate<T, N, (N <= 16 ? 8 : 4)>(G, b1, w, v, x == x0);
      }
      if (two) {
        const StridedPencil<T, N> b2(const_cast<cpx<T>*>(jb.tab2) +
                                     (static_cast<size_t>(x) * jb.tabM2 + a.rows2[jb.dir_off]) * n + o);
        pencil_accumulate<T, N, (N <= 16 ? 8 : 4)>(G2, b2, w, v, x == x0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Packed isotropic register-resident gain (FP32, N = 16, 24).  For real
// f_p, f_q the fields t3 = IFFT(E1 F_p) and t4 = IFFT(E2 F_q) are real except
// for what their Nyquist planes contribute: with phi the phase of E and
// delta(k) = phi(k) + phi(-k) (zero unless a component of k is N/2, where
// lfreq maps -k back to -N/2),
//   Im t = IFFT(D),  D(k) = F(k) E(k) (1 - exp(-i delta(k))) / (2i),
// which lives on the three Nyquist planes.  Each plane's part is a 2D
// inverse transform broadcast along the third axis with the sign (-1)^x,
// (-1)^y or (-1)^z (owner order x, then y, then z, so no k is counted
// twice):  Im t(x, y, z) = (-1)^x W0(y, z) + (-1)^y W1(x, z) + (-1)^z W2(x, y).
// One complex transform P = IFFT(E1 F_p + i E2 F_q) = t3 + i t4 then gives
//   Re t3 = Re P + Im t4,   Re t4 = Im P - Im t3,
// so a direction product costs one 3D inverse transform and six N x N 2D
// transforms instead of two 3D inverse transforms, with the same t3 t4 (exact
// in exact arithmetic).  The sum over directions, the forward transform per
// radial node and the G^ accumulation are those of k_vss_gain_rr1_iso.
// ---------------------------------------------------------------------------
// the packed isotropic kernel is the automatic choice once it is measured faster
constexpr bool kVssPackDefault = false;

template <int N>
struct VssPackShape {
  static constexpr int park = (N == 24) ? 4 : 0;  // accumulator values kept in the spare LDS
};

template <typename T, int N>
__global__ __launch_bounds__(VssRR1Shape<N>::NT) void k_vss_gain_rr1_isop(VssGainArgs<T> a) {
  constexpr int NT = VssRR1Shape<N>::NT, PP = VssRR1Shape<N>::PP, NN = N * N, n = NN * N;
  constexpr int KS = VssRR1Shape<N>::KS, h = N / 2, KP = VssPackShape<N>::park;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X = reinterpret_cast<cpx<T>*>(smem_raw);
  __shared__ cpx<T> e1[3][N], e2[3][N];
  __shared__ cpx<T> cs[2][8];       // (1 - exp(-i delta_S)) / (2i), S the Nyquist components (bit 0 x, 1 y, 2 z)
  constexpr int PW = N + 1;         // padded row pitch of the W planes (spreads the row reads over the banks)
  __shared__ cpx<T> Wp[6][N * PW];  // W0, W1, W2 of t3, then of t4
  __shared__ cpx<T> apark[KP > 0 ? KP : 1][NT];
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  cpx<T>* Xp = X + g * KS;  // kx plane g
  const int per_pt = (a.nitems / (a.npts * a.ngroups)) * a.ngroups;
  cpx<T> acc[N], v[N];
  for (int item = xcd_slot(); item < a.nitems; item += gridDim.x) {
    const int pt = item / per_pt, rem = item - pt * per_pt;
    const int j = rem / a.ngroups, grp = rem - j * a.ngroups;
    const int x0 = grp * a.ppg, x1 = min(x0 + a.ppg, a.Nrho);
    const VssJob<T>& jb = a.job[j];
    const cpx<T>* F1 = a.F + (static_cast<size_t>(jb.sp) * a.npts + pt) * n;
    const cpx<T>* F2 = a.F + (static_cast<size_t>(jb.sq) * a.npts + pt) * n;
    const double mu1 = jb.mu1, mu2 = jb.mu2;
    const size_t o = static_cast<size_t>(g) * NN + l;  // this thread's G^ pencil (kx = g, kz = l), stride N
    const bool two = jb.out2 >= 0;
    const StridedPencil<T, N> G(a.Ghat + ((static_cast<size_t>(jb.out1) * a.npts + pt) * a.ngroups + grp) * n + o);
    const StridedPencil<T, N> G2(a.Ghat + ((static_cast<size_t>(two ? jb.out2 : jb.out1) * a.npts + pt) *
                                               a.ngroups + grp) * n + o);
    for (int x = x0; x < x1; ++x) {
      const double alpha = a.alpha[x];
#pragma unroll
      for (int k = 0; k < N - KP; ++k) acc[k] = cpx<T>{T(0), T(0)};
#pragma unroll
      for (int k = 0; k < KP; ++k) apark[k][tid] = cpx<T>{T(0), T(0)};  // thread-private slots
      for (int iy = 0; iy < jb.ndir; ++iy) {
        const int y = a.dirs[jb.dir_off + iy];
        __syncthreads();  // readers of the tables, W planes and LDS field of the previous product are done
        if (tid < 3 * N) {
          const int d = tid / N, k = tid - d * N;
          const double th = alpha * a.sz[y * 3 + d] * lfreq(k, N);
          double sn, c0;
          sincos(mu1 * th, &sn, &c0);
          e1[d][k] = cpx<T>{static_cast<T>(c0), static_cast<T>(sn)};
          sincos(mu2 * th, &sn, &c0);
          e2[d][k] = cpx<T>{static_cast<T>(c0), static_cast<T>(-sn)};
        } else if (tid < 3 * N + 16) {
          const int f = (tid - 3 * N) >> 3, S = (tid - 3 * N) & 7;
          double sum = 0.0;
          for (int d = 0; d < 3; ++d)
            if (S & (1 << d)) sum += a.sz[y * 3 + d];
          // delta = phi(k) + phi(-k): -mu1 alpha N sum for t3, +mu2 alpha N sum for t4
          const double delta = (f == 0 ? -mu1 : mu2) * alpha * N * sum;
          double sn, c0;
          sincos(delta, &sn, &c0);
          cs[f][S] = cpx<T>{static_cast<T>(0.5 * sn), static_cast<T>(-0.5 * (1.0 - c0))};
        }
        __syncthreads();
        // D on the six Nyquist planes (plane q = 3 f + owner, element (i, jj)),
        // every thread a few elements so that the F loads overlap
        for (int e = tid; e < 6 * NN; e += NT) {
          const int q = e / NN, ij = e - q * NN, i = ij / N, jj = ij - i * N, f = q / 3, ow = q - 3 * f;
          int kx, ky, kz;
          bool zero;
          if (ow == 0) { kx = h; ky = i; kz = jj; zero = false; }
          else if (ow == 1) { kx = i; ky = h; kz = jj; zero = (i == h); }
          else { kx = i; ky = jj; kz = h; zero = (i == h) || (jj == h); }
          const int S = (kx == h ? 1 : 0) | (ky == h ? 2 : 0) | (kz == h ? 4 : 0);
          const cpx<T>* Fs = f ? F2 : F1;
          const cpx<T> (*ef)[N] = f ? e2 : e1;
          Wp[q][i * PW + jj] = zero ? cpx<T>{T(0), T(0)}
                           : Fs[kx * NN + ky * N + kz] * ((ef[0][kx] * ef[1][ky] * ef[2][kz]) * cs[f][S]);
        }
        __syncthreads();
        // W planes, rows
        if (tid < 6 * N) {
          const int q = tid / N, i = tid - q * N;
#pragma unroll
          for (int jj = 0; jj < N; ++jj) v[jj] = Wp[q][i * PW + jj];
          FFT<T, N, +1, true>::run(v);
#pragma unroll
          for (int jj = 0; jj < N; ++jj) Wp[q][i * PW + jj] = v[jj];
        }
        __syncthreads();
        // W planes, columns
        if (tid < 6 * N) {
          const int q = tid / N, jj = tid - q * N;
#pragma unroll
          for (int i = 0; i < N; ++i) v[i] = Wp[q][i * PW + jj];
          FFT<T, N, +1, true>::run(v);
#pragma unroll
          for (int i = 0; i < N; ++i) Wp[q][i * PW + jj] = v[i];
        }
        // packed y-pencil (kx = g, kz = l): E1 F_p + i E2 F_q
        {
          const StridedPencil<T, N> s1(const_cast<cpx<T>*>(F1) + g * NN + l);
          const StridedPencil<T, N> s2(const_cast<cpx<T>*>(F2) + g * NN + l);
          const cpx<T> ea = e1[0][g] * e1[2][l], eb = e2[0][g] * e2[2][l];
#pragma unroll
          for (int ky = 0; ky < N; ++ky) v[ky] = s1[ky] * (ea * e1[1][ky]) + mul_i(s2[ky] * (eb * e2[1][ky]));
        }
        vss_rr1_inverse_tail<T, N>(v, X, Xp, g, l);  // P at (x, y = l, z = g); the tail's barriers order the W columns too
        {
          const T sl = (l & 1) ? T(-1) : T(1), sg = (g & 1) ? T(-1) : T(1);
          const T w30 = Wp[0][l * PW + g].x, w40 = Wp[3][l * PW + g].x;
#pragma unroll
          for (int xx = 0; xx < N; ++xx) {
            const T sx = (xx & 1) ? T(-1) : T(1);
            const T i3 = sx * w30 + sl * Wp[1][xx * PW + g].x + sg * Wp[2][xx * PW + l].x;
            const T i4 = sx * w40 + sl * Wp[4][xx * PW + g].x + sg * Wp[5][xx * PW + l].x;
            const cpx<T> t3{v[xx].x + i4, i3}, t4{v[xx].y - i3, i4};
            if (xx < N - KP) acc[xx] = acc[xx] + t3 * t4;
            else apark[xx - (N - KP)][tid] = apark[xx - (N - KP)][tid] + t3 * t4;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < KP; ++k) acc[N - KP + k] = apark[k][tid];
      // G^ += w b(x) FFT(sum_y t3 t4)
      FFT<T, N, -1, true>::run(acc);  // x: (kx, y = l, z = g)
      __syncthreads();
#pragma unroll
      for (int kx = 0; kx < N; ++kx) X[kx * KS + g * N + l] = acc[kx];
      __syncthreads();
#pragma unroll
      for (int z = 0; z < N; ++z) v[z] = X[g * KS + z * N + l];
      FFT<T, N, -1, true>::run(v);  // z: (kx = g, y = l, kz)
      __syncthreads();
#pragma unroll
      for (int kz = 0; kz < N; ++kz) Xp[kz * PP + l] = v[kz];
      __syncthreads();
#pragma unroll
      for (int yy = 0; yy < N; ++yy) v[yy] = Xp[l * PP + yy];
      FFT<T, N, -1, true>::run(v);  // y: (kx = g, ky, kz = l)
      const T w = static_cast<T>(a.wgt[x]);
      {
        const StridedPencil<T, N> b1(const_cast<cpx<T>*>(jb.tab1) +
                                     (static_cast<size_t>(x) * jb.tabM1 + a.rows1[jb.dir_off]) * n + o);
        pencil_accumulate<T, N, (N <= 16 ? 8 : 4)>(G, b1, w, v, x == x0);
      }
      if (two) {
        const StridedPencil<T, N> b2(const_cast<cpx<T>*>(jb.tab2) +
                                     (static_cast<size_t>(x) * jb.tabM2 + a.rows2[jb.dir_off]) * n + o);
        pencil_accumulate<T, N, (N <= 16 ? 8 : 4)>(G2, b2, w, v, x == x0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// finalize: one item = (species p, point)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Slab gain for isotropic scattering (eta = 0 for every pair), the FP64 form
// at Nv 16 / 24.  A workgroup owns a (point, job, radial group, slab) item;
// slab s holds the XS = N/NS planes x = s + NS j of the two inverse
// transforms t3 = IFFT(E1 F_p) and t4 = IFFT(E2 F_q) in LDS, and its two
// halves transform one each.  Per direction product:
//   x-pass  pencils (ky, kz) of the half's spectrum times its x phase
//           factors, folded over kx = r + XS m with exp(2 pi i m s / NS), the
//           yz phase and the twiddle exp(2 pi i r s / N) applied after the
//           fold, an XS-point inverse FFT: plane j, row ky, column kz of the
//           half's slab;
//   z-pass  one row (plane j, ky) per thread;
//   y-pass  one column (plane j, z) per thread, back into its own column;
//   product S_x[y] += t3 t4 (complex), the t3 half's thread of column (j, z)
//           taking y < N/2 and the t4 half's the rest, in registers over the
//           job's directions.
// At the end of radial node x the slab of S_x is forward transformed along
// y and z into T[job][point][x]; k_vss_xfwd completes the transform along x
// and forms G^ = sum_x w_x b(x) FFT(S_x) for both outputs of the job.  Same
// jobs, tables and sums as k_vss_gain<T, N, true>; only the order of the
// transforms' stages differs.
// ---------------------------------------------------------------------------
template <int N, typename T = double>
struct VssSlabShape {
  // slabs per point: the two FP64 slabs fill the LDS at NS = 3 for 24^3
  // (FP32: NS = 2, 576 threads); 16^3 fits whole
  static constexpr int NS = (N == 24) ? (sizeof(T) == 4 ? 2 : 3) : 1;
  static constexpr int XS = N / NS;
  static constexpr int NH = XS * N;  // threads per slab: one row of it each
  static constexpr int NT = 2 * NH;  // the t3 half and the t4 half of the workgroup
  // plane rows padded to N + 1 (row and column accesses conflict free with
  // immediate offsets); swizzled rows made the compiler hold one LDS offset
  // per element across the item loop and spill (MI355X, 24^3: FP64 100.7 ->
  // 84.1 ms, FP32 with NS = 2 50.1 -> 42.5 ms per 2048 mixture points)
  static constexpr int RP = N + 1;
  static constexpr int PL = N * RP;          // complex per plane
  static constexpr int planes = XS * PL;     // complex per slab (storage)
  static constexpr int cells = XS * N * N;   // values per slab
  // offset of row r, column c within a plane
  static __device__ __forceinline__ int at(int r, int c) { return r * RP + c; }
  static_assert(NS * XS == N, "slabs must tile the x axis");
  static_assert((N * N) % NH == 0, "x-pass pencils split evenly over a half");
  static_assert(N % 2 == 0, "the product's y range splits in halves");
};

template <typename T>
struct VssSlabArgs {
  VssGainArgs<T> g;  // spectra, jobs, tables, phase inputs, item split
  cpx<T>* Tb;        // [njobs][npts][Nrho][n]
  int njobs;
};

template <typename T, int N, int S>
__device__ __forceinline__ void vss_slab_x_pass(const cpx<T>* __restrict__ F, cpx<T>* X, const cpx<T> (*e)[N],
                                                int th) {
  using SS = VssSlabShape<N, T>;
  constexpr int NS = SS::NS, XS = SS::XS, NH = SS::NH, NN = N * N;
#pragma unroll 1
  for (int i = 0; i < NN / NH; ++i) {
    const int idx = th + NH * i;  // ky * N + kz
    const int ky = idx / N, kz = idx - (idx / N) * N;
    const cpx<T> eyz = e[1][ky] * e[2][kz];
    // the XS loads of fold term m + 1 are in flight while term m is
    // accumulated (the loads of the whole pencil held at once spill at two
    // waves per SIMD); the x phase factors are read from LDS per pencil
    // through an opaque offset, so they are not held in registers
    int eo = 0;
    asm volatile("" : "+v"(eo));
    const cpx<T>* ex = &e[0][0] + eo;
    const cpx<T>* P = F + idx;
    cpx<T> U[XS], ld[2][XS];
#pragma unroll
    for (int r = 0; r < XS; ++r) {
      U[r] = cpx<T>{T(0), T(0)};
      ld[0][r] = P[r * NN];
    }
    // (the 576-thread FP32 slab, NS = 2, loads one term at a time: its
    // prefetch measured slower, 44.6 against 42.3 ms per 2048 mixture points
    // at 24^3 on MI355X)
    constexpr bool PF = SS::NT <= 512;
    static_for<0, NS>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      if constexpr (PF && m + 1 < NS) {
#pragma unroll
        for (int r = 0; r < XS; ++r) ld[(m + 1) & 1][r] = P[(r + XS * (m + 1)) * NN];
      }
      if constexpr (!PF && m > 0) {
#pragma unroll
        for (int r = 0; r < XS; ++r) ld[m & 1][r] = P[(r + XS * m) * NN];
      }
      static_for<0, XS>([&](auto rc) {
        constexpr int r = decltype(rc)::value, kx = r + XS * m;
        U[r] = U[r] + FFT<T, NS, +1>::template tw<(m * S) % NS>(ld[m & 1][r] * ex[kx]);
      });
      if constexpr (!PF) __builtin_amdgcn_sched_barrier(0);
    });
    static_for<0, XS>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      U[r] = FFT<T, N, +1>::template tw<(r * S) % N>(U[r] * eyz);
    });
    FFT<T, XS, +1, true>::run(U);
#pragma unroll
    for (int j = 0; j < XS; ++j) X[j * SS::PL + SS::at(ky, kz)] = U[j];
  }
}

// x-pass of this half's spectrum (t3 half: F_p; t4 half: F_q); th is the
// thread's index within its half
template <typename T, int N>
__device__ __forceinline__ void vss_slab_x_dispatch(int s, const cpx<T>* F, cpx<T>* X, const cpx<T> (*e)[N],
                                                    int th) {
  constexpr int NS = VssSlabShape<N, T>::NS;
  if constexpr (NS == 1) {
    vss_slab_x_pass<T, N, 0>(F, X, e, th);
  } else if constexpr (NS == 2) {
    if (s == 0) vss_slab_x_pass<T, N, 0>(F, X, e, th);
    else vss_slab_x_pass<T, N, 1>(F, X, e, th);
  } else {
    static_assert(NS == 3, "slab counts 1, 2 and 3");
    if (s == 0) vss_slab_x_pass<T, N, 0>(F, X, e, th);
    else if (s == 1) vss_slab_x_pass<T, N, 1>(F, X, e, th);
    else vss_slab_x_pass<T, N, 2>(F, X, e, th);
  }
}

template <typename T, int N>
__global__ __launch_bounds__((VssSlabShape<N, T>::NT)) void k_vss_gain_slab(VssSlabArgs<T> sa) {
  using SS = VssSlabShape<N, T>;
  constexpr int NS = SS::NS, NH = SS::NH, NN = N * N, n = NN * N, HY = N / 2;
  const VssGainArgs<T>& a = sa.g;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X1 = reinterpret_cast<cpx<T>*>(smem_raw);  // t3 slab
  cpx<T>* X2 = X1 + SS::planes;                       // t4 slab
  __shared__ cpx<T> eph[2][2][3][N];  // [buffer][t3, t4][axis][k]: this product's phases and the next's
  const int tid = threadIdx.x;
  // half h transforms t3 (h = 0) or t4 (h = 1); within it, thread th has
  // plane j and row / column c
  const int h = tid / NH, th = tid - (tid / NH) * NH;
  const int j = th / N, c = th - (th / N) * N;
  cpx<T>* Xh = h ? X2 : X1;
  // items run point-major, so one point's spectra serve its consecutive
  // items (jobs, groups, slabs) from one XCD's L2
  const int per_pt = sa.njobs * a.ngroups * NS;
  cpx<T> sacc[HY];  // S_x at y = h * N/2 + i of column (j, c)
  for (int item = xcd_slot(); item < a.nitems; item += gridDim.x) {
    const int pt = item / per_pt, rem = item - pt * per_pt;
    const int jb_i = rem / (a.ngroups * NS), rem2 = rem - jb_i * (a.ngroups * NS);
    const int grp = rem2 / NS, s = rem2 - grp * NS;
    const int x0 = grp * a.ppg, x1 = min(x0 + a.ppg, a.Nrho);
    const VssJob<T>& jb = a.job[jb_i];
    const cpx<T>* Fh = a.F + (static_cast<size_t>(h ? jb.sq : jb.sp) * a.npts + pt) * n;
    const double mu1 = jb.mu1, mu2 = jb.mu2;
    // phase factors of direction product (x, iy) into buffer b
    auto phases = [&](int b, int x, int iy) {
      if (tid < 3 * N) {
        const int d = tid / N, k = tid - d * N;
        const int y = a.dirs[jb.dir_off + iy];
        const double th_ = a.alpha[x] * a.sz[y * 3 + d] * lfreq(k, N);
        double sn, cs;
        sincos(mu1 * th_, &sn, &cs);
        eph[b][0][d][k] = cpx<T>{static_cast<T>(cs), static_cast<T>(sn)};
        sincos(mu2 * th_, &sn, &cs);
        eph[b][1][d][k] = cpx<T>{static_cast<T>(cs), static_cast<T>(-sn)};
      }
    };
    __syncthreads();  // the previous item's readers of eph are done
    if (x0 < x1 && jb.ndir > 0) phases(0, x0, 0);
    int it = 0;  // products of this item: eph[it & 1] holds the current phases
    for (int x = x0; x < x1; ++x) {
#pragma unroll
      for (int i = 0; i < HY; ++i) sacc[i] = cpx<T>{T(0), T(0)};
      for (int iy = 0; iy < jb.ndir; ++iy, ++it) {
        __syncthreads();  // the previous product's readers of X1, X2 are done; eph[it & 1] is in place
        vss_slab_x_dispatch<T, N>(s, Fh, Xh, eph[it & 1][h], th);
        __syncthreads();
        {  // z-pass: row (j, ky = c) of this half's slab
          cpx<T>* pl = Xh + j * SS::PL;
          cpx<T> v[N];
#pragma unroll
          for (int b = 0; b < N; ++b) v[b] = pl[SS::at(c, b)];
          FFT<T, N, +1, true>::run(v);
#pragma unroll
          for (int b = 0; b < N; ++b) pl[SS::at(c, b)] = v[b];
        }
        __syncthreads();
        // the next product's phase factors into the other buffer (the x-pass
        // above was this product's last reader of eph)
        {
          const int iyn = (iy + 1 < jb.ndir) ? iy + 1 : 0, xn = (iy + 1 < jb.ndir) ? x : x + 1;
          if (xn < x1) phases((it + 1) & 1, xn, iyn);
        }
        {  // y-pass: column (j, z = c) of this half's slab, back into its own column
          cpx<T>* pl = Xh + j * SS::PL;
          cpx<T> v[N];
#pragma unroll
          for (int ky = 0; ky < N; ++ky) v[ky] = pl[SS::at(ky, c)];
          FFT<T, N, +1, true>::run(v);
#pragma unroll
          for (int y = 0; y < N; ++y) pl[SS::at(y, c)] = v[y];
        }
        __syncthreads();
        {  // S_x += t3 t4 on this thread's half of column (j, c)
          const cpx<T>* p1 = X1 + j * SS::PL;
          const cpx<T>* p2 = X2 + j * SS::PL;
#pragma unroll
          for (int i = 0; i < HY; ++i) {
            const int y = h * HY + i, o = SS::at(y, c);
            sacc[i] = sacc[i] + p1[o] * p2[o];
          }
        }
      }
      // forward transform of the slab of S_x along y and z (the t3 half's
      // threads; the halves first gather each column's S in LDS)
      __syncthreads();  // every product read of X1, X2 is done
      {
        cpx<T>* pl = X1 + j * SS::PL;
#pragma unroll
        for (int i = 0; i < HY; ++i) {
          const int y = h * HY + i;
          pl[SS::at(y, c)] = sacc[i];
        }
      }
      __syncthreads();
      if (h == 0) {
        cpx<T>* pl = X1 + j * SS::PL;
        cpx<T> v[N];
#pragma unroll
        for (int y = 0; y < N; ++y) v[y] = pl[SS::at(y, c)];
        FFT<T, N, -1, true>::run(v);
#pragma unroll
        for (int ky = 0; ky < N; ++ky) pl[SS::at(ky, c)] = v[ky];
      }
      __syncthreads();
      if (h == 0) {
        cpx<T>* pl = X1 + j * SS::PL;
        cpx<T> v[N];
#pragma unroll
        for (int b = 0; b < N; ++b) v[b] = pl[SS::at(c, b)];
        FFT<T, N, -1, true>::run(v);
#pragma unroll
        for (int b = 0; b < N; ++b) pl[SS::at(c, b)] = v[b];
      }
      __syncthreads();
      {
        cpx<T>* Tp = sa.Tb + ((static_cast<size_t>(jb_i) * a.npts + pt) * a.Nrho + x) * n;
        for (int e = tid; e < SS::cells; e += SS::NT) {
          const int jj = e / NN, r = e - jj * NN, ky = r / N, kz = r - ky * N;
          Tp[static_cast<size_t>(s + NS * jj) * NN + r] = X1[jj * SS::PL + SS::at(ky, kz)];
        }
      }
    }
  }
}

// x-forward transform of the slab kernel's T and G^_out = sum_x w_x b(x)
// FFT(S_x) for both outputs of a job (tables read at the job's first
// direction row, as the isotropic multi-pass kernel).  A block of 256 threads
// takes 128 (job, point, pencil (ky, kz)) columns; waves 2i and 2i + 1 share
// 64 of them and keep the lower and the upper half of the kx sums.
template <typename T, int N, int K0>
__device__ __forceinline__ void vss_xfwd_half(const VssSlabArgs<T>& sa, int jb_i, int pt, int idx) {
  constexpr int NN = N * N, n = NN * N, H = N / 2;
  const VssGainArgs<T>& a = sa.g;
  const VssJob<T>& jb = a.job[jb_i];
  const bool two = jb.out2 >= 0;
  const int row1 = a.rows1[jb.dir_off], row2 = two ? a.rows2[jb.dir_off] : 0;
  cpx<T> acc1[H], acc2[H];
#pragma unroll
  for (int k = 0; k < H; ++k) acc1[k] = acc2[k] = cpx<T>{T(0), T(0)};
  for (int x = 0; x < a.Nrho; ++x) {
    const cpx<T>* Tp = sa.Tb + ((static_cast<size_t>(jb_i) * a.npts + pt) * a.Nrho + x) * n + idx;
    cpx<T> v[N];
#pragma unroll
    for (int xx = 0; xx < N; ++xx) v[xx] = Tp[static_cast<size_t>(xx) * NN];
    FFT<T, N, -1>::run(v);
    const T w = static_cast<T>(a.wgt[x]);
    const cpx<T>* b1 = jb.tab1 + (static_cast<size_t>(x) * jb.tabM1 + row1) * n + idx;
#pragma unroll
    for (int k = 0; k < H; ++k) acc1[k] = acc1[k] + (w * b1[static_cast<size_t>(k + K0) * NN]) * v[k + K0];
    if (two) {
      const cpx<T>* b2 = jb.tab2 + (static_cast<size_t>(x) * jb.tabM2 + row2) * n + idx;
#pragma unroll
      for (int k = 0; k < H; ++k) acc2[k] = acc2[k] + (w * b2[static_cast<size_t>(k + K0) * NN]) * v[k + K0];
    }
  }
  cpx<T>* G1 = a.Ghat + (static_cast<size_t>(jb.out1) * a.npts + pt) * n + idx;
#pragma unroll
  for (int k = 0; k < H; ++k) G1[static_cast<size_t>(k + K0) * NN] = acc1[k];
  if (two) {
    cpx<T>* G2 = a.Ghat + (static_cast<size_t>(jb.out2) * a.npts + pt) * n + idx;
#pragma unroll
    for (int k = 0; k < H; ++k) G2[static_cast<size_t>(k + K0) * NN] = acc2[k];
  }
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_vss_xfwd(VssSlabArgs<T> sa) {
  constexpr int NN = N * N;
  const int w = threadIdx.x / 64, lane = threadIdx.x - (threadIdx.x / 64) * 64;
  const long col = static_cast<long>(blockIdx.x) * 128 + (w >> 1) * 64 + lane;
  const long ncol = static_cast<long>(sa.njobs) * sa.g.npts * NN;
  if (col >= ncol) return;
  const long jp = col / NN;
  const int idx = static_cast<int>(col - jp * NN);
  const int jb_i = static_cast<int>(jp / sa.g.npts), pt = static_cast<int>(jp - static_cast<long>(jb_i) * sa.g.npts);
  if (w & 1)
    vss_xfwd_half<T, N, N / 2>(sa, jb_i, pt, idx);
  else
    vss_xfwd_half<T, N, 0>(sa, jb_i, pt, idx);
}

template <typename T>
struct VssFinArgs {
  const cpx<T>* Ghat;  // [npairs][npts][ngroups][n]
  const cpx<T>* F;     // [nspcs][npts][n]
  const double* f[kVssMaxSpecies];  // [npts][n] per species
  double* Q[kVssMaxSpecies];        // [npts][n] per species
  cpx<T>* Z;           // scratch [nslots][n]
  const cpx<T>* c[kVssMaxPairs];
  double prefac[kVssMaxPairs];
  int nspcs, npts, ngroups;
};

template <typename T, int N>
__global__ __launch_bounds__(Shape<N>::NT) void k_vss_finalize(VssFinArgs<T> a) {
  constexpr int NT = Shape<N>::NT, PB = Shape<N>::PB, NN = N * N, n = N * N * N;
  __shared__ cpx<T> lds[PB * N * (N + 1)];
  cpx<T>* Z = a.Z + static_cast<size_t>(blockIdx.x) * n;
  const int nitems = a.nspcs * a.npts;
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int p = item / a.npts, pt = item - p * a.npts;
    for (int kx0 = 0; kx0 < N; kx0 += PB) {
      const int cnt = min(PB, N - kx0);
      slab_batch<T, N, NT, +1>(
          lds, cnt,
          [&](int j, int ky, int kz) -> cpx<T> {
            const int kx = kx0 + j;
            const int o = kx * NN + ky * N + kz;
            const int om = ((N - kx) % N) * NN + ((N - ky) % N) * N + (N - kz) % N;
            cpx<T> acc{T(0), T(0)};
            for (int q = 0; q < a.nspcs; ++q) {
              const int pair = p * a.nspcs + q;
              const cpx<T>* G = a.Ghat + (static_cast<size_t>(pair) * a.npts + pt) * a.ngroups * n;
              cpx<T> g0 = G[o], g1 = G[om];
              for (int gg = 1; gg < a.ngroups; ++gg) {
                g0 = g0 + G[static_cast<size_t>(gg) * n + o];
                g1 = g1 + G[static_cast<size_t>(gg) * n + om];
              }
              const cpx<T>* Fq = a.F + (static_cast<size_t>(q) * a.npts + pt) * n;
              const cpx<T>* cq = a.c[pair];
              const cpx<T> l0 = cq[o] * Fq[o], l1 = cq[om] * Fq[om];
              const cpx<T> gh = T(0.5) * (g0 + conj(g1)), lh = T(0.5) * (l0 + conj(l1));
              acc = acc + static_cast<T>(a.prefac[pair]) * (gh + mul_i(lh));
            }
            return acc;
          },
          [&](int j, int r, int c, cpx<T> v) { Z[(kx0 + j) * NN + r * N + c] = v; });
    }
    const double* f = a.f[p] + static_cast<size_t>(pt) * n;
    double* Q = a.Q[p] + static_cast<size_t>(pt) * n;
    pencil_pass<T, N, NT, +1>([&](int kx, int idx) { return Z[kx * NN + idx]; },
                              [&](int idx, cpx<T>* v) {
                                // f's pencil loaded before Q is stored (loads in flight together)
                                double fv[N];
#pragma unroll
                                for (int x = 0; x < N; ++x) fv[x] = f[x * NN + idx];
#pragma unroll
                                for (int x = 0; x < N; ++x)
                                  Q[x * NN + idx] = static_cast<double>(v[x].x) - fv[x] * static_cast<double>(v[x].y);
                              });
  }
}

// ---------------------------------------------------------------------------
// FP64 projection onto the mixture's collision invariants: the mass of every
// species, total momentum sum_s m_s int c Q_s and total energy
// sum_s m_s int |c|^2 Q_s.  Minimum-norm correction over the stacked vector
// (Q_0, ..., Q_{S-1}): d = -C^T (C C^T)^{-1} C Q, C_j(s, r) = phi_j(s, r) cw.
// ---------------------------------------------------------------------------
constexpr int kVssMaxCon = kVssMaxSpecies + 4;

struct VssProjArgs {
  double* Q[kVssMaxSpecies];  // [npts][n] per species
  double Ainv[kVssMaxCon * kVssMaxCon];
  double m[kVssMaxSpecies];
  double L, cw;
  int N, nspcs, npts;
};

__global__ __launch_bounds__(256) void k_vss_project(VssProjArgs a) {
  __shared__ double red[kVssMaxCon][256];
  __shared__ double lam[kVssMaxCon];
  const int n = a.N * a.N * a.N, tid = threadIdx.x, S = a.nspcs, nc = S + 4;
  const double hh = 2.0 * a.L / a.N, c0 = -a.L + 0.5 * hh;
  for (int pt = blockIdx.x; pt < a.npts; pt += gridDim.x) {
    // red rows: mass of species 0..S-1, then momentum x, y, z and energy
    double mass[kVssMaxSpecies] = {};
    double m1 = 0, m2 = 0, m3 = 0, m4 = 0;
#pragma unroll
    for (int s = 0; s < kVssMaxSpecies; ++s) {
      if (s < S) {
        const double* Q = a.Q[s] + static_cast<size_t>(pt) * n;
        const double ms = a.m[s];
        for (int r = tid; r < n; r += 256) {
          const int I = r / (a.N * a.N), J = (r / a.N) % a.N, K = r % a.N;
          const double cx = c0 + I * hh, cy = c0 + J * hh, cz = c0 + K * hh, q = Q[r];
          mass[s] += q;
          m1 += ms * q * cx;
          m2 += ms * q * cy;
          m3 += ms * q * cz;
          m4 += ms * q * (cx * cx + cy * cy + cz * cz);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < kVssMaxSpecies; ++s)
      if (s < S) red[s][tid] = mass[s] * a.cw;
    red[S][tid] = m1 * a.cw;
    red[S + 1][tid] = m2 * a.cw;
    red[S + 2][tid] = m3 * a.cw;
    red[S + 3][tid] = m4 * a.cw;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (tid < st)
        for (int j = 0; j < nc; ++j) red[j][tid] += red[j][tid + st];
      __syncthreads();
    }
    if (tid < nc) {
      double v = 0;
      for (int k = 0; k < nc; ++k) v += a.Ainv[tid * nc + k] * red[k][0];
      lam[tid] = v;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kVssMaxSpecies; ++s) {
      if (s < S) {
        double* Q = a.Q[s] + static_cast<size_t>(pt) * n;
        const double ms = a.m[s], ls = lam[s];
        for (int r = tid; r < n; r += 256) {
          const int I = r / (a.N * a.N), J = (r / a.N) % a.N, K = r % a.N;
          const double cx = c0 + I * hh, cy = c0 + J * hh, cz = c0 + K * hh;
          Q[r] -= a.cw * (ls + ms * (lam[S] * cx + lam[S + 1] * cy + lam[S + 2] * cz +
                                     lam[S + 3] * (cx * cx + cy * cy + cz * cz)));
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
struct VssHandle {
  int magic = 0x76737331;
  int N = 0, n = 0, Nrho = 0, M = 0, Mpre = 0, nspcs = 0, npairs = 0, prec = 64, project = 0;
  double L = 0, sw = 0, cw = 0;
  size_t tsize = 8;
  double masses[kVssMaxSpecies] = {};
  double prefac[kVssMaxPairs] = {};
  int sp[kVssMaxPairs] = {}, sq[kVssMaxPairs] = {};
  double mu1[kVssMaxPairs] = {}, mu2[kVssMaxPairs] = {};
  void* b[kVssMaxPairs] = {};
  void* c[kVssMaxPairs] = {};
  // antipodal sharing: merged self-pair tables [Nrho][M/2][n] and the job lists
  void* bsym[kVssMaxSpecies] = {};
  int sym_ok = 0, sym = 0;
  std::vector<int> anti, half;
  struct Job {
    int sp, sq, out1, out2, tab1, tab2, sym1, tabM1, tabM2, ndir, dir_off;
  };
  std::vector<Job> jobs_plain, jobs_sym;
  int* dirs_d = nullptr;
  int* rows1_d = nullptr;
  int* rows2_d = nullptr;
  double* alpha = nullptr;
  double* wgt = nullptr;
  double* szd = nullptr;
  double Ainv[kVssMaxCon * kVssMaxCon] = {};
  // workspace
  size_t cap_pts = 0, cap_items = 0;
  int cap_slots = 0;
  double* fc = nullptr;  // [nspcs][cap_pts][n]
  double* Qc = nullptr;
  void* F = nullptr;     // [nspcs][cap_pts][n] cpx<T>
  void* Ghat = nullptr;  // [cap_items][n] cpx<T>
  void* Z1 = nullptr;
  void* Z2 = nullptr;
  void* Z3 = nullptr;
  int groups_override = 0;
  int slots_override = 0;  // cap on the persistent gain grid (0: occupancy)
  int algorithm = 0;       // 0: automatic, 1: multi-pass, 4: register-resident (FP32, Nv 16 / 24),
                           // 6: packed isotropic register-resident, 7: slab (isotropic, Nv 16 / 24)
  void* Tb = nullptr;      // slab kernel's forward-transformed slabs [njobs][cap_pts][Nrho][n]
  size_t cap_tb = 0;
  int iso = 0;             // eta = 0 for every pair: b_pq(x, y) independent of y
  int last_algorithm = 0;
  int last_groups = 0, last_slots = 0, threads = 0;
  int timing = 0;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double acc_ms[5] = {0, 0, 0, 0, 0};
  int acc_n[5] = {0, 0, 0, 0, 0};

  ~VssHandle() {
    for (int k = 0; k < kVssMaxPairs; ++k) {
      (void)hipFree(b[k]);
      (void)hipFree(c[k]);
    }
    for (int k = 0; k < kVssMaxSpecies; ++k) (void)hipFree(bsym[k]);
    (void)hipFree(dirs_d); (void)hipFree(rows1_d); (void)hipFree(rows2_d);
    (void)hipFree(alpha); (void)hipFree(wgt); (void)hipFree(szd);
    (void)hipFree(fc); (void)hipFree(Qc); (void)hipFree(F); (void)hipFree(Ghat);
    (void)hipFree(Z1); (void)hipFree(Z2); (void)hipFree(Z3); (void)hipFree(Tb);
    for (auto& pr : pending) {
      (void)hipEventDestroy(pr.second.first);
      (void)hipEventDestroy(pr.second.second);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

static int invert_n(const double* A, double* Ainv, int m) {
  std::vector<double> w(static_cast<size_t>(m) * 2 * m);
  auto at = [&](int i, int j) -> double& { return w[static_cast<size_t>(i) * 2 * m + j]; };
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < 2 * m; ++j) at(i, j) = (j < m) ? A[i * m + j] : (j - m == i ? 1.0 : 0.0);
  for (int col = 0; col < m; ++col) {
    int piv = col;
    for (int r = col + 1; r < m; ++r)
      if (std::fabs(at(r, col)) > std::fabs(at(piv, col))) piv = r;
    if (std::fabs(at(piv, col)) < 1e-300) return 1;
    for (int j = 0; j < 2 * m; ++j) std::swap(at(col, j), at(piv, j));
    const double d = at(col, col);
    for (int j = 0; j < 2 * m; ++j) at(col, j) /= d;
    for (int r = 0; r < m; ++r)
      if (r != col) {
        const double fct = at(r, col);
        for (int j = 0; j < 2 * m; ++j) at(r, j) -= fct * at(col, j);
      }
  }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) Ainv[i * m + j] = at(i, j + m);
  return 0;
}

template <typename T, int N>
static int vss_slots() {
  int o1 = 1, o2 = 1, o3 = 1;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o1, reinterpret_cast<const void*>(k_vss_gain<T, N, false>),
                                                     Shape<N>::NT, 0);
  int o4 = 1;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o4, reinterpret_cast<const void*>(k_vss_gain<T, N, true>),
                                                     Shape<N>::NT, 0);
  o1 = std::max(o1, o4);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o2, reinterpret_cast<const void*>(k_vhs_prep<T, N>),
                                                     Shape<N>::NT, 0);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o3, reinterpret_cast<const void*>(k_vss_finalize<T, N>),
                                                     Shape<N>::NT, 0);
  return std::max(1, std::min(o1, std::min(o2, o3))) * cu_count();
}

template <typename T, int N>
static int vss_ensure_ws(VssHandle* h, int pts) {
  constexpr int n = N * N * N;
  if (h->cap_slots == 0) {
    const int slots = vss_slots<T, N>();
    DGFS_HIP_TRY(hipMalloc(&h->Z1, static_cast<size_t>(slots) * n * sizeof(cpx<T>)));
    DGFS_HIP_TRY(hipMalloc(&h->Z2, static_cast<size_t>(slots) * n * sizeof(cpx<T>)));
    if (h->iso) DGFS_HIP_TRY(hipMalloc(&h->Z3, static_cast<size_t>(slots) * n * sizeof(cpx<T>)));
    h->cap_slots = slots;
  }
  if (static_cast<size_t>(pts) > h->cap_pts) {
    (void)hipFree(h->F); (void)hipFree(h->fc); (void)hipFree(h->Qc);
    h->F = nullptr; h->fc = h->Qc = nullptr;
    h->cap_pts = 0;
    const size_t cnt = static_cast<size_t>(h->nspcs) * pts * n;
    DGFS_HIP_TRY(hipMalloc(&h->F, cnt * sizeof(cpx<T>)));
    DGFS_HIP_TRY(hipMalloc(&h->fc, cnt * sizeof(double)));
    DGFS_HIP_TRY(hipMalloc(&h->Qc, cnt * sizeof(double)));
    h->cap_pts = pts;
  }
  return DGFS_OK;
}

// points per launch: per point the species' f, Q (FP64) and spectra, plus the
// gain partials of every pair (worst case one per radial node); under 48 GiB
static int vss_chunk_points(const VssHandle* h, int total) {
  // (the slab kernel's T, one transform per job and radial node, is bounded by
  // the same npairs * Nrho transforms)
  const double per = static_cast<double>(h->n) *
                     (h->nspcs * (16.0 + 2.0 * h->tsize) + h->npairs * 2.0 * h->tsize * h->Nrho * (h->iso ? 2 : 1));
  const double budget = 48.0 * 1024.0 * 1024.0 * 1024.0;
  const int cap = std::max(1, static_cast<int>(budget / per));
  if (total <= cap) return std::max(total, 1);
  const int nchunks = (total + cap - 1) / cap;
  return (total + nchunks - 1) / nchunks;
}

// gain algorithm of this call: 4 (register-resident, FP32 at Nv 16 / 24)
// where it is built and not switched off; in FP64 with isotropic scattering
// at Nv 16 / 24 the slab kernel (7: 16.1 against 67.0 ms per 2048 mixture
// points at 16^3, 192.3 against 301.2 ms at 24^3 on MI355X); else 1
// (multi-pass)
template <typename T, int N>
static int vss_rr_kind(const VssHandle* h) {
  constexpr bool rr_ok = std::is_same<T, float>::value && (N == 16 || N == 24);
  constexpr bool slab_ok = (N == 16 || N == 24);
  if (h->algorithm == 1) return 1;
  if (h->algorithm == 7) return (slab_ok && h->iso) ? 7 : 1;
  if (rr_ok) {
    // FP32 isotropic: the slab kernel (per 2048 mixture points on MI355X,
    // 16^3: 10.6 against 12.0 ms for the one-pencil kernel; 24^3: 42.5
    // against 48.9 ms).  Algorithm 6 is validated to FP32 at 16 / 24 when set.
    if (h->iso && h->algorithm == 0) return 7;
    return 4;
  }
  // FP64 isotropic: the slab kernel
  if (slab_ok && h->iso && h->algorithm == 0) return 7;
  return 1;
}

// fs[s], Qs[s]: [npts][n] contiguous per species
template <typename T, int N>
static int vss_run_chunk(VssHandle* h, const double* const* fs, double* const* Qs, int npts, hipStream_t st) {
  constexpr int n = N * N * N;
  const int slots = h->slots_override > 0 ? std::min(h->cap_slots, h->slots_override) : h->cap_slots;
  const std::vector<VssHandle::Job>& jobs = h->sym ? h->jobs_sym : h->jobs_plain;
  const int njobs = static_cast<int>(jobs.size());
  const int kind = vss_rr_kind<T, N>(h);
  int slab_ns = 1;
  if constexpr (N == 16 || N == 24) slab_ns = VssSlabShape<N, T>::NS;
  int groups = h->groups_override;
  if (groups <= 0) {
    if (kind == 7) {  // (point, job, group, slab) items, one workgroup per CU
      const int base = njobs * std::max(npts, 1) * slab_ns;
      groups = std::max(1, std::min(h->Nrho, (cu_count() + base - 1) / base));
    } else {
      const int base = njobs * std::max(npts, 1);
      groups = std::max(1, std::min(h->Nrho, (2 * slots + base - 1) / base));
    }
  }
  const int ppg = (h->Nrho + groups - 1) / groups;
  groups = (h->Nrho + ppg - 1) / ppg;
  const size_t items = static_cast<size_t>(njobs) * npts * groups;
  // the slab kernel's x-forward sum leaves one G^ per (pair, point)
  const int fin_groups = (kind == 7) ? 1 : groups;
  const size_t outs = static_cast<size_t>(h->npairs) * npts * fin_groups;
  if (outs > h->cap_items) {
    (void)hipFree(h->Ghat);
    h->Ghat = nullptr;
    h->cap_items = 0;
    DGFS_HIP_TRY(hipMalloc(&h->Ghat, outs * n * sizeof(cpx<T>)));
    h->cap_items = outs;
  }
  h->last_groups = groups;
  h->last_slots = slots;
  h->threads = Shape<N>::NT;

  cpx<T>* F = static_cast<cpx<T>*>(h->F);
  {
    KTimerT<VssHandle> t0(h, 0, st);
    for (int s = 0; s < h->nspcs; ++s) {
      PrepArgs<T> pa{fs[s], F + static_cast<size_t>(s) * npts * n, nullptr,
                     static_cast<cpx<T>*>(h->Z1), npts};
      hipLaunchKernelGGL((k_vhs_prep<T, N>), dim3(std::min(npts, slots)), dim3(Shape<N>::NT), 0, st, pa);
      DGFS_LAUNCH_CHECK("vss_prep");
    }
    t0.stop();
  }

  VssGainArgs<T> ga{};
  ga.F = F;
  ga.Ghat = static_cast<cpx<T>*>(h->Ghat);
  ga.Z1 = static_cast<cpx<T>*>(h->Z1);
  ga.Z2 = static_cast<cpx<T>*>(h->Z2);
  ga.Z3 = static_cast<cpx<T>*>(h->Z3);
  ga.iso = h->iso;
  ga.alpha = h->alpha;
  ga.wgt = h->wgt;
  ga.sz = h->szd;
  ga.dirs = h->dirs_d;
  ga.rows1 = h->rows1_d;
  ga.rows2 = h->rows2_d;
  for (int k = 0; k < njobs; ++k) {
    const VssHandle::Job& jb = jobs[k];
    VssJob<T>& o = ga.job[k];
    o.tab1 = static_cast<const cpx<T>*>(jb.sym1 ? h->bsym[jb.tab1] : h->b[jb.tab1]);
    o.tab2 = (jb.out2 >= 0) ? static_cast<const cpx<T>*>(h->b[jb.tab2]) : nullptr;
    o.mu1 = h->masses[jb.sq] / (h->masses[jb.sp] + h->masses[jb.sq]);
    o.mu2 = h->masses[jb.sp] / (h->masses[jb.sp] + h->masses[jb.sq]);
    o.sp = jb.sp;
    o.sq = jb.sq;
    o.out1 = jb.out1;
    o.out2 = jb.out2;
    o.tabM1 = jb.tabM1;
    o.tabM2 = jb.tabM2;
    o.ndir = jb.ndir;
    o.dir_off = jb.dir_off;
  }
  ga.nitems = static_cast<int>(items);
  ga.npts = npts;
  ga.ngroups = groups;
  ga.ppg = ppg;
  ga.Nrho = h->Nrho;
  if (kind == 7) {
    if constexpr (N == 16 || N == 24) {
      using SS = VssSlabShape<N, T>;
      const size_t tbytes = static_cast<size_t>(njobs) * npts * h->Nrho * n * sizeof(cpx<T>);
      if (tbytes > h->cap_tb) {
        (void)hipFree(h->Tb);
        h->Tb = nullptr;
        h->cap_tb = 0;
        DGFS_HIP_TRY(hipMalloc(&h->Tb, tbytes));
        h->cap_tb = tbytes;
      }
      VssSlabArgs<T> sa{ga, static_cast<cpx<T>*>(h->Tb), njobs};
      sa.g.nitems = static_cast<int>(items * SS::NS);
      const size_t lds = 2 * static_cast<size_t>(SS::planes) * sizeof(cpx<T>);
      const void* kfn = reinterpret_cast<const void*>(k_vss_gain_slab<T, N>);
      DGFS_HIP_TRY(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      int occ = 1;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, SS::NT, lds) != hipSuccess || occ < 1) occ = 1;
      int grid = static_cast<int>(std::min<size_t>(items * SS::NS, static_cast<size_t>(occ) * cu_count()));
      if (h->slots_override > 0) grid = std::min(grid, h->slots_override);
      KTimerT<VssHandle> t1(h, 1, st);
      hipLaunchKernelGGL((k_vss_gain_slab<T, N>), dim3(grid), dim3(SS::NT), lds, st, sa);
      DGFS_LAUNCH_CHECK("vss_gain_slab");
      const long cols = static_cast<long>(njobs) * npts * N * N;
      hipLaunchKernelGGL((k_vss_xfwd<T, N>), dim3(static_cast<unsigned>((cols + 127) / 128)), dim3(256), 0, st, sa);
      DGFS_LAUNCH_CHECK("vss_xfwd");
      t1.stop();
      h->last_slots = grid;
      h->threads = SS::NT;
      h->last_algorithm = 7;
    }
  } else if (kind == 4) {
    if constexpr (std::is_same<T, float>::value && (N == 16 || N == 24)) {
      const size_t lds = static_cast<size_t>(VssRR1Shape<N>::field) * sizeof(cpx<T>);
      const bool pack = h->iso && (h->algorithm == 6 || (h->algorithm == 0 && kVssPackDefault));
      const void* kfn = pack ? reinterpret_cast<const void*>(k_vss_gain_rr1_isop<T, N>)
                        : h->iso ? reinterpret_cast<const void*>(k_vss_gain_rr1_iso<T, N>)
                                 : reinterpret_cast<const void*>(k_vss_gain_rr1<T, N>);
      DGFS_HIP_TRY(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      int occ = 1;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, VssRR1Shape<N>::NT, lds) != hipSuccess || occ < 1)
        occ = 1;
      int grid = static_cast<int>(std::min<size_t>(items, static_cast<size_t>(occ) * cu_count()));
      if (h->slots_override > 0) grid = std::min(grid, h->slots_override);
      KTimerT<VssHandle> t1(h, 1, st);
      if (grid > h->cap_slots) grid = h->cap_slots;  // the isotropic kernel parks t3 in a scratch cube per workgroup
      if (pack)
        hipLaunchKernelGGL((k_vss_gain_rr1_isop<T, N>), dim3(grid), dim3(VssRR1Shape<N>::NT), lds, st, ga);
      else if (h->iso)
        hipLaunchKernelGGL((k_vss_gain_rr1_iso<T, N>), dim3(grid), dim3(VssRR1Shape<N>::NT), lds, st, ga);
      else
        hipLaunchKernelGGL((k_vss_gain_rr1<T, N>), dim3(grid), dim3(VssRR1Shape<N>::NT), lds, st, ga);
      DGFS_LAUNCH_CHECK("vss_gain_rr1");
      t1.stop();
      h->last_slots = grid;
      h->threads = VssRR1Shape<N>::NT;
      h->last_algorithm = pack ? 6 : h->iso ? 5 : 4;
    }
  } else {
    KTimerT<VssHandle> t1(h, 1, st);
    if (h->iso)
      hipLaunchKernelGGL((k_vss_gain<T, N, true>), dim3(static_cast<int>(std::min<size_t>(items, slots))),
                         dim3(Shape<N>::NT), 0, st, ga);
    else
      hipLaunchKernelGGL((k_vss_gain<T, N, false>), dim3(static_cast<int>(std::min<size_t>(items, slots))),
                         dim3(Shape<N>::NT), 0, st, ga);
    DGFS_LAUNCH_CHECK("vss_gain");
    t1.stop();
    h->last_algorithm = 1;
  }

  VssFinArgs<T> fa{};
  fa.Ghat = static_cast<const cpx<T>*>(h->Ghat);
  fa.F = F;
  for (int sp = 0; sp < h->nspcs; ++sp) {
    fa.f[sp] = fs[sp];
    fa.Q[sp] = Qs[sp];
  }
  fa.Z = static_cast<cpx<T>*>(h->Z1);
  for (int k = 0; k < h->npairs; ++k) {
    fa.c[k] = static_cast<const cpx<T>*>(h->c[k]);
    fa.prefac[k] = h->prefac[k];
  }
  fa.nspcs = h->nspcs;
  fa.npts = npts;
  fa.ngroups = fin_groups;
  {
    KTimerT<VssHandle> t2(h, 2, st);
    hipLaunchKernelGGL((k_vss_finalize<T, N>), dim3(std::min(h->nspcs * npts, slots)), dim3(Shape<N>::NT), 0,
                       st, fa);
    DGFS_LAUNCH_CHECK("vss_finalize");
    t2.stop();
  }

  if (h->project) {
    VssProjArgs pr{};
    for (int sp = 0; sp < h->nspcs; ++sp) pr.Q[sp] = Qs[sp];
    std::memcpy(pr.Ainv, h->Ainv, sizeof(pr.Ainv));
    for (int s = 0; s < h->nspcs; ++s) pr.m[s] = h->masses[s];
    pr.L = h->L;
    pr.cw = h->cw;
    pr.N = N;
    pr.nspcs = h->nspcs;
    pr.npts = npts;
    KTimerT<VssHandle> t3(h, 3, st);
    hipLaunchKernelGGL(k_vss_project, dim3(std::min(npts, 4 * cu_count())), dim3(256), 0, st, pr);
    DGFS_LAUNCH_CHECK("vss_project");
    t3.stop();
  }
  return DGFS_OK;
}

// f[s], Q[s]: per-species registers (AoSoA) or contiguous [npts][n] arrays
template <typename T, int N>
static int vss_eval_impl(VssHandle* h, const double* const* f, double* const* Q, bool aosoa, int nupts,
                         int ldim, int vsize, int e0, int ne, int npts, hipStream_t st) {
  constexpr int n = N * N * N;
  const int S = h->nspcs;
  if (aosoa) {
    const int total = nupts * ne;
    if (total == 0) return DGFS_OK;
    const int per_chunk_elems = std::max(1, vss_chunk_points(h, total) / std::max(nupts, 1));
    for (int c0 = 0; c0 < ne; c0 += per_chunk_elems) {
      const int cne = std::min(per_chunk_elems, ne - c0);
      const int pts = cne * nupts;
      int rc = vss_ensure_ws<T, N>(h, pts);
      if (rc) return rc;
      for (int s = 0; s < S; ++s) {
        rc = launch_aosoa_gather(f[s], h->fc + static_cast<size_t>(s) * pts * n, nupts, ldim, vsize, e0 + c0,
                                 cne, st);
        if (rc) return rc;
      }
      const double* fs[kVssMaxSpecies];
      double* Qs[kVssMaxSpecies];
      for (int s = 0; s < S; ++s) {
        fs[s] = h->fc + static_cast<size_t>(s) * pts * n;
        Qs[s] = h->Qc + static_cast<size_t>(s) * pts * n;
      }
      rc = vss_run_chunk<T, N>(h, fs, Qs, pts, st);
      if (rc) return rc;
      for (int s = 0; s < S; ++s) {
        rc = launch_aosoa_scatter(h->Qc + static_cast<size_t>(s) * pts * n, Q[s], nupts, ldim, vsize, e0 + c0,
                                  cne, st);
        if (rc) return rc;
      }
    }
    return DGFS_OK;
  }
  const int chunk = vss_chunk_points(h, npts);
  for (int p0 = 0; p0 < npts; p0 += chunk) {
    const int pts = std::min(chunk, npts - p0);
    int rc = vss_ensure_ws<T, N>(h, pts);
    if (rc) return rc;
    const double* fs[kVssMaxSpecies];
    double* Qs[kVssMaxSpecies];
    for (int s = 0; s < S; ++s) {
      fs[s] = f[s] + static_cast<size_t>(p0) * n;
      Qs[s] = Q[s] + static_cast<size_t>(p0) * n;
    }
    rc = vss_run_chunk<T, N>(h, fs, Qs, pts, st);
    if (rc) return rc;
  }
  return DGFS_OK;
}

template <typename T>
static int vss_dispatch(VssHandle* h, const double* const* f, double* const* Q, bool aosoa, int nupts, int ldim,
                        int vsize, int e0, int ne, int npts, hipStream_t st) {
  switch (h->N) {
#define DGFS_VSS_CASE(NV) \
  case NV:                \
    return vss_eval_impl<T, NV>(h, f, Q, aosoa, nupts, ldim, vsize, e0, ne, npts, st);
    DGFS_VSS_CASE(4) DGFS_VSS_CASE(6) DGFS_VSS_CASE(8) DGFS_VSS_CASE(12) DGFS_VSS_CASE(16) DGFS_VSS_CASE(24)
    DGFS_VSS_CASE(32)
#undef DGFS_VSS_CASE
    default:
      return fail(DGFS_ERR_UNSUPPORTED, "Nv=%d is not built", h->N);
  }
}

static VssHandle* as_vss(void* p) {
  auto* h = static_cast<VssHandle*>(p);
  if (!h || h->magic != 0x76737331) return nullptr;
  return h;
}

template <typename T>
static int vss_build_tables(VssHandle* h, const dgfs_vss_desc* d) {
  const int n = h->n, N = h->N, K = h->Nrho * h->M, Mpre = d->Mpre;
  const double swpre = 4.0 * 3.14159265358979323846 / Mpre;
  double *qz = nullptr, *qw = nullptr, *szpre = nullptr, *Byj = nullptr, *Bsum = nullptr;
  struct Freer {
    double** ptrs[5];
    ~Freer() {
      for (auto p : ptrs) (void)hipFree(*p);
    }
  } freer{{&qz, &qw, &szpre, &Byj, &Bsum}};
  DGFS_HIP_TRY(hipMalloc(&qz, h->Nrho * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&qw, h->Nrho * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&szpre, static_cast<size_t>(Mpre) * 3 * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&Byj, static_cast<size_t>(h->M) * Mpre * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&Bsum, static_cast<size_t>(Mpre) * sizeof(double)));
  DGFS_HIP_TRY(hipMemcpy(qz, d->qz, h->Nrho * sizeof(double), hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(qw, d->qw, h->Nrho * sizeof(double), hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(szpre, d->szpre, static_cast<size_t>(Mpre) * 3 * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> B(static_cast<size_t>(h->M) * Mpre), Bs(Mpre);
  for (int k = 0; k < h->npairs; ++k) {
    const int p = h->sp[k], q = h->sq[k];
    const double eta = d->eta[p * h->nspcs + q], gamma = d->gamma[p * h->nspcs + q];
    std::fill(Bs.begin(), Bs.end(), 0.0);
    for (int y = 0; y < h->M; ++y)
      for (int j = 0; j < Mpre; ++j) {
        const double dot = d->sz[3 * y] * d->szpre[3 * j] + d->sz[3 * y + 1] * d->szpre[3 * j + 1] +
                           d->sz[3 * y + 2] * d->szpre[3 * j + 2];
        // (1 + s.s')^eta; the base is clamped at 0 where rounding of an
        // antipodal pair makes it a tiny negative number
        const double v = std::pow(std::max(0.0, 1.0 + dot), eta);
        B[static_cast<size_t>(y) * Mpre + j] = v;
        Bs[j] += v;
      }
    DGFS_HIP_TRY(hipMemcpy(Byj, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice));
    DGFS_HIP_TRY(hipMemcpy(Bsum, Bs.data(), Bs.size() * sizeof(double), hipMemcpyHostToDevice));
    DGFS_HIP_TRY(hipMalloc(&h->b[k], static_cast<size_t>(K) * n * sizeof(cpx<T>)));
    DGFS_HIP_TRY(hipMalloc(&h->c[k], static_cast<size_t>(n) * sizeof(cpx<T>)));
    const double mpq = h->masses[q] / (h->masses[p] + h->masses[q]);
    hipLaunchKernelGGL(k_vss_btab<T>, dim3((n + 255) / 256, K), dim3(256), 0, 0, static_cast<cpx<T>*>(h->b[k]),
                       Byj, szpre, qz, h->L, gamma, mpq, swpre, N, h->M, Mpre);
    DGFS_LAUNCH_CHECK("vss_btab");
    hipLaunchKernelGGL(k_vss_ctab<T>, dim3((n + 255) / 256), dim3(256), 0, 0, static_cast<cpx<T>*>(h->c[k]), Bsum,
                       szpre, qz, qw, h->L, gamma, h->sw, swpre, N, h->Nrho, Mpre);
    DGFS_LAUNCH_CHECK("vss_ctab");
    // the host arrays are reused for the next pair
    DGFS_HIP_TRY(hipDeviceSynchronize());
  }
  if (h->sym_ok) {
    // b(x, y) + b(x, -y) for the self pairs: the sum of the two rows of B
    // taken in FP64 before the table is formed
    const int Mh = static_cast<int>(h->half.size());
    std::vector<double> Bh(static_cast<size_t>(Mh) * Mpre);
    for (int p = 0; p < h->nspcs; ++p) {
      const double eta = d->eta[p * h->nspcs + p], gamma = d->gamma[p * h->nspcs + p];
      for (int i = 0; i < Mh; ++i)
        for (int j = 0; j < Mpre; ++j) {
          double v = 0.0;
          for (const int y : {h->half[i], h->anti[h->half[i]]}) {
            const double dot = d->sz[3 * y] * d->szpre[3 * j] + d->sz[3 * y + 1] * d->szpre[3 * j + 1] +
                               d->sz[3 * y + 2] * d->szpre[3 * j + 2];
            v += std::pow(std::max(0.0, 1.0 + dot), eta);
          }
          Bh[static_cast<size_t>(i) * Mpre + j] = v;
        }
      DGFS_HIP_TRY(hipMemcpy(Byj, Bh.data(), Bh.size() * sizeof(double), hipMemcpyHostToDevice));
      DGFS_HIP_TRY(hipMalloc(&h->bsym[p], static_cast<size_t>(h->Nrho) * Mh * n * sizeof(cpx<T>)));
      hipLaunchKernelGGL(k_vss_btab<T>, dim3((n + 255) / 256, h->Nrho * Mh), dim3(256), 0, 0,
                         static_cast<cpx<T>*>(h->bsym[p]), Byj, szpre, qz, h->L, gamma, 0.5, swpre, N, Mh, Mpre);
      DGFS_LAUNCH_CHECK("vss_btab_sym");
      DGFS_HIP_TRY(hipDeviceSynchronize());
    }
  }
  return DGFS_OK;
}

// antipodes of the design (exact negation) and the two job lists
static int vss_build_jobs(VssHandle* h, const double* sz) {
  const int M = h->M, S = h->nspcs;
  h->anti.assign(M, -1);
  h->sym_ok = 1;
  for (int y = 0; y < M && h->sym_ok; ++y) {
    for (int z = 0; z < M; ++z)
      if (z != y && sz[3 * z] == -sz[3 * y] && sz[3 * z + 1] == -sz[3 * y + 1] && sz[3 * z + 2] == -sz[3 * y + 2]) {
        h->anti[y] = z;
        break;
      }
    if (h->anti[y] < 0) h->sym_ok = 0;
  }
  h->half.clear();
  if (h->sym_ok)
    for (int y = 0; y < M; ++y)
      if (h->anti[y] > y) h->half.push_back(y);
  if (h->sym_ok && 2 * static_cast<int>(h->half.size()) != M) h->sym_ok = 0;
  std::vector<int> dirs, rows1, rows2;
  auto add = [&](std::vector<VssHandle::Job>& list, VssHandle::Job jb, const std::vector<int>& d,
                 const std::vector<int>& r1, const std::vector<int>& r2) {
    jb.ndir = static_cast<int>(d.size());
    jb.dir_off = static_cast<int>(dirs.size());
    dirs.insert(dirs.end(), d.begin(), d.end());
    rows1.insert(rows1.end(), r1.begin(), r1.end());
    rows2.insert(rows2.end(), r2.begin(), r2.end());
    list.push_back(jb);
  };
  std::vector<int> all(M);
  for (int y = 0; y < M; ++y) all[y] = y;
  for (int p = 0; p < S; ++p)
    for (int q = 0; q < S; ++q) {
      const int k = p * S + q;
      add(h->jobs_plain, VssHandle::Job{p, q, k, -1, k, 0, 0, M, 0, 0, 0}, all, all, all);
    }
  if (h->sym_ok) {
    // cross pairs first (a full sweep of directions each), then the self pairs
    for (int p = 0; p < S; ++p)
      for (int q = p + 1; q < S; ++q)
        add(h->jobs_sym, VssHandle::Job{p, q, p * S + q, q * S + p, p * S + q, q * S + p, 0, M, M, 0, 0}, all,
            all, h->anti);
    std::vector<int> hrow(h->half.size());
    for (size_t i = 0; i < hrow.size(); ++i) hrow[i] = static_cast<int>(i);
    for (int p = 0; p < S; ++p)
      add(h->jobs_sym, VssHandle::Job{p, p, p * S + p, -1, p, 0, 1, static_cast<int>(hrow.size()), 0, 0, 0},
          h->half, hrow, hrow);
  }
  h->sym = h->sym_ok;
  const size_t bytes = dirs.size() * sizeof(int);
  DGFS_HIP_TRY(hipMalloc(&h->dirs_d, bytes));
  DGFS_HIP_TRY(hipMalloc(&h->rows1_d, bytes));
  DGFS_HIP_TRY(hipMalloc(&h->rows2_d, bytes));
  DGFS_HIP_TRY(hipMemcpy(h->dirs_d, dirs.data(), bytes, hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(h->rows1_d, rows1.data(), bytes, hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(h->rows2_d, rows2.data(), bytes, hipMemcpyHostToDevice));
  return DGFS_OK;
}

}  // namespace dgfs

using namespace dgfs;

extern "C" int dgfs_vss_create(const dgfs_vss_desc* d, void** out) {
  if (!d || !out) return fail(DGFS_ERR_INVALID_VALUE, "null descriptor or handle pointer");
  *out = nullptr;
  switch (d->Nv) {
    case 4: case 6: case 8: case 12: case 16: case 24: case 32: break;
    default: return fail(DGFS_ERR_UNSUPPORTED, "Nv=%d unsupported (built: 4 6 8 12 16 24 32)", d->Nv);
  }
  if (d->nspcs < 1 || d->nspcs > kVssMaxSpecies)
    return fail(DGFS_ERR_INVALID_VALUE, "nspcs must be 1..%d (got %d)", kVssMaxSpecies, d->nspcs);
  if (d->Nrho < 1 || d->M < 1 || d->Mpre < 1)
    return fail(DGFS_ERR_INVALID_VALUE, "Nrho, M and Mpre must be >= 1");
  if (!d->qz || !d->qw || !d->sz || !d->szpre || !d->masses || !d->gamma || !d->eta || !d->prefac)
    return fail(DGFS_ERR_INVALID_VALUE, "null table pointer in the descriptor");
  if (d->precision != 64 && d->precision != 32)
    return fail(DGFS_ERR_INVALID_VALUE, "precision must be 64 or 32 (got %d)", d->precision);
  if (!(d->L > 0) || !(d->cw > 0) || !(d->sw > 0)) return fail(DGFS_ERR_INVALID_VALUE, "L, cw and sw must be positive");
  for (int s = 0; s < d->nspcs; ++s)
    if (!(d->masses[s] > 0)) return fail(DGFS_ERR_INVALID_VALUE, "species masses must be positive");

  auto h = std::make_unique<VssHandle>();
  h->N = d->Nv;
  h->n = d->Nv * d->Nv * d->Nv;
  h->Nrho = d->Nrho;
  h->M = d->M;
  h->Mpre = d->Mpre;
  h->nspcs = d->nspcs;
  h->npairs = d->nspcs * d->nspcs;
  h->prec = d->precision;
  h->project = d->project ? 1 : 0;
  h->L = d->L;
  h->sw = d->sw;
  h->cw = d->cw;
  h->tsize = (d->precision == 64) ? 8 : 4;
  for (int s = 0; s < h->nspcs; ++s) h->masses[s] = d->masses[s];
  for (int p = 0; p < h->nspcs; ++p)
    for (int q = 0; q < h->nspcs; ++q) {
      const int k = p * h->nspcs + q;
      h->sp[k] = p;
      h->sq[k] = q;
      h->prefac[k] = d->prefac[k];
      h->mu1[k] = h->masses[q] / (h->masses[p] + h->masses[q]);
      h->mu2[k] = h->masses[p] / (h->masses[p] + h->masses[q]);
    }
  const double pi = 3.14159265358979323846;
  std::vector<double> alpha(h->Nrho), wgt(h->Nrho);
  for (int x = 0; x < h->Nrho; ++x) {
    alpha[x] = pi / h->L * d->qz[x];
    wgt[x] = d->qw[x] * h->sw / static_cast<double>(h->n);
  }
  DGFS_HIP_TRY(hipMalloc(&h->alpha, h->Nrho * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&h->wgt, h->Nrho * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&h->szd, h->M * 3 * sizeof(double)));
  DGFS_HIP_TRY(hipMemcpy(h->alpha, alpha.data(), h->Nrho * sizeof(double), hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(h->wgt, wgt.data(), h->Nrho * sizeof(double), hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(h->szd, d->sz, h->M * 3 * sizeof(double), hipMemcpyHostToDevice));
  h->iso = 1;
  for (int k = 0; k < h->npairs; ++k)
    if (d->eta[k] != 0.0) h->iso = 0;
  int rc = vss_build_jobs(h.get(), d->sz);
  if (rc) return rc;
  rc = (h->prec == 64) ? vss_build_tables<double>(h.get(), d) : vss_build_tables<float>(h.get(), d);
  if (rc) return rc;

  // Gram matrix of the invariants: rows mass_s (s < S), then m_s c_x, m_s c_y,
  // m_s c_z and m_s |c|^2 over all species blocks; A_jk = sum phi_j phi_k cw^2
  const int S = h->nspcs, nc = S + 4, N = h->N;
  std::vector<double> A(static_cast<size_t>(nc) * nc, 0.0);
  const double hh = 2.0 * h->L / N, c0 = -h->L + 0.5 * hh;
  double mom[5][5] = {};
  for (int r = 0; r < h->n; ++r) {
    const int I = r / (N * N), J = (r / N) % N, K = r % N;
    const double cx = c0 + I * hh, cy = c0 + J * hh, cz = c0 + K * hh;
    const double phi[5] = {1.0, cx, cy, cz, cx * cx + cy * cy + cz * cz};
    for (int j = 0; j < 5; ++j)
      for (int k = 0; k < 5; ++k) mom[j][k] += phi[j] * phi[k];
  }
  const double w2 = h->cw * h->cw;
  for (int s = 0; s < S; ++s) {
    const double m = h->masses[s];
    A[s * nc + s] += mom[0][0] * w2;
    for (int j = 0; j < 4; ++j) {
      A[s * nc + S + j] += m * mom[0][1 + j] * w2;
      A[(S + j) * nc + s] += m * mom[0][1 + j] * w2;
      for (int k = 0; k < 4; ++k) A[(S + j) * nc + S + k] += m * m * mom[1 + j][1 + k] * w2;
    }
  }
  if (invert_n(A.data(), h->Ainv, nc)) return fail(DGFS_ERR_INVALID_VALUE, "singular invariant Gram matrix");
  *out = h.release();
  return DGFS_OK;
}

extern "C" int dgfs_vss_destroy(void* p) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  (void)hipDeviceSynchronize();
  resolve_timing(h);
  h->magic = 0;
  delete h;
  return DGFS_OK;
}

static int vss_check_ptrs(const VssHandle* h, const double* const* f, double* const* Q) {
  if (!f || !Q) return fail(DGFS_ERR_INVALID_VALUE, "null species pointer array");
  for (int s = 0; s < h->nspcs; ++s) {
    if (!f[s] || !Q[s]) return fail(DGFS_ERR_INVALID_VALUE, "null distribution pointer of species %d", s);
    for (int t = 0; t < h->nspcs; ++t)
      if (static_cast<const void*>(f[s]) == static_cast<const void*>(Q[t]))
        return fail(DGFS_ERR_INVALID_VALUE, "f and Q may not alias (species %d / %d)", s, t);
  }
  return DGFS_OK;
}

extern "C" int dgfs_vss_eval(void* p, const double* const* f, double* const* Q, int nupts, int ldim, int vsize,
                             int neles, int e0, int ne, void* stream) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  int rc = vss_check_ptrs(h, f, Q);
  if (rc) return rc;
  if (vsize != h->n) return fail(DGFS_ERR_INVALID_VALUE, "vsize %d != Nv^3 = %d", vsize, h->n);
  if (nupts < 1 || neles < 0 || e0 < 0 || ne < 0 || e0 + ne > neles)
    return fail(DGFS_ERR_INVALID_VALUE, "bad element range e0=%d ne=%d neles=%d nupts=%d", e0, ne, neles, nupts);
  const long need = static_cast<long>((neles + 31) / 32) * vsize * 32;
  if (ldim < need) return fail(DGFS_ERR_INVALID_VALUE, "ldim %d < %ld (AoSoA row of %d elements)", ldim, need, neles);
  auto st = static_cast<hipStream_t>(stream);
  return (h->prec == 64) ? vss_dispatch<double>(h, f, Q, true, nupts, ldim, vsize, e0, ne, 0, st)
                         : vss_dispatch<float>(h, f, Q, true, nupts, ldim, vsize, e0, ne, 0, st);
}

extern "C" int dgfs_vss_eval_contig(void* p, const double* const* f, double* const* Q, int npts, void* stream) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  int rc = vss_check_ptrs(h, f, Q);
  if (rc) return rc;
  if (npts < 0) return fail(DGFS_ERR_INVALID_VALUE, "npts must be >= 0");
  if (npts == 0) return DGFS_OK;
  auto st = static_cast<hipStream_t>(stream);
  return (h->prec == 64) ? vss_dispatch<double>(h, f, Q, false, 0, 0, h->n, 0, 0, npts, st)
                         : vss_dispatch<float>(h, f, Q, false, 0, 0, h->n, 0, 0, npts, st);
}

extern "C" int dgfs_vss_get_config(void* p, int* ngroups, int* nslots, int* threads) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (ngroups) *ngroups = h->last_groups;
  if (nslots) *nslots = h->last_slots;
  if (threads) *threads = h->threads;
  return DGFS_OK;
}

extern "C" int dgfs_vss_set_groups(void* p, int ngroups) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (ngroups < 0 || ngroups > h->Nrho) return fail(DGFS_ERR_INVALID_VALUE, "ngroups must be 0..Nrho");
  h->groups_override = ngroups;
  return DGFS_OK;
}

extern "C" int dgfs_vss_set_algorithm(void* p, int algorithm) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (algorithm != 0 && algorithm != 1 && algorithm != 4 && algorithm != 6 && algorithm != 7)
    return fail(DGFS_ERR_INVALID_VALUE,
                "algorithm must be 0 (automatic), 1 (multi-pass), 4 (register-resident), 6 (packed isotropic) or "
                "7 (slab)");
  if ((algorithm == 4 || algorithm == 6) && !(h->prec == 32 && (h->N == 16 || h->N == 24)))
    return fail(DGFS_ERR_UNSUPPORTED, "the register-resident VSS gain is built for FP32 at Nv 16 and 24");
  if (algorithm == 6 && !h->iso)
    return fail(DGFS_ERR_UNSUPPORTED, "the packed VSS gain needs isotropic scattering (eta = 0 for every pair)");
  if (algorithm == 7 && !(h->iso && (h->N == 16 || h->N == 24)))
    return fail(DGFS_ERR_UNSUPPORTED, "the slab VSS gain is built for isotropic scattering at Nv 16 and 24");
  h->algorithm = algorithm;
  return DGFS_OK;
}

extern "C" int dgfs_vss_get_algorithm(void* p, int* algorithm) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (!algorithm) return fail(DGFS_ERR_INVALID_VALUE, "null output");
  *algorithm = h->last_algorithm;
  return DGFS_OK;
}

extern "C" int dgfs_vss_set_slots(void* p, int nslots) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (nslots < 0) return fail(DGFS_ERR_INVALID_VALUE, "nslots must be >= 0");
  h->slots_override = nslots;
  return DGFS_OK;
}

extern "C" int dgfs_vss_set_symmetry(void* p, int on) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (on && !h->sym_ok)
    return fail(DGFS_ERR_UNSUPPORTED, "the direction design does not hold the antipode of every direction");
  h->sym = on ? 1 : 0;
  return DGFS_OK;
}

extern "C" int dgfs_vss_get_symmetry(void* p, int* on) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (!on) return fail(DGFS_ERR_INVALID_VALUE, "null output");
  *on = h->sym;
  return DGFS_OK;
}

extern "C" int dgfs_vss_set_timing(void* p, int on) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  resolve_timing(h);
  h->timing = on ? 1 : 0;
  for (int k = 0; k < 5; ++k) {
    h->acc_ms[k] = 0.0;
    h->acc_n[k] = 0;
  }
  return DGFS_OK;
}

extern "C" int dgfs_vss_get_timing(void* p, double* ms, int* counts) {
  VssHandle* h = as_vss(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VSS handle");
  if (!ms || !counts) return fail(DGFS_ERR_INVALID_VALUE, "null output");
  resolve_timing(h);
  for (int k = 0; k < 5; ++k) {
    ms[k] = h->acc_ms[k];
    counts[k] = h->acc_n[k];
  }
  return DGFS_OK;
}
