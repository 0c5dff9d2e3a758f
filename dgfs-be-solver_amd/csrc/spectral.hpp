// Device building blocks shared by the collision operators (single-species
// VHS and multi-species VSS): LDS-staged 2D slab transforms, register pencil
// transforms, the spectrum prep kernel, AoSoA gather/scatter launchers and the
// per-kernel event timing the handles expose.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>
#include <vector>

#include "common.hpp"
#include "cplx.hpp"

namespace dgfs {

// ---------------------------------------------------------------------------
// workgroup building blocks
// ---------------------------------------------------------------------------

// 2D FFTs of `cnt` N x N planes staged through LDS (row stride N+1 to spread
// banks).  load(j, r, c) gives element (r, c) of plane j; the plane is
// transformed along c then along r; store(j, r, c, v) receives the result.
template <typename T, int N, int NT, int SIGN, class Load, class Store>
__device__ __forceinline__ void slab_batch(cpx<T>* lds, int cnt, Load load, Store store) {
  constexpr int NP = N + 1, PL = N * NP, NN = N * N;
  const int tid = threadIdx.x;
  for (int e = tid; e < cnt * NN; e += NT) {
    const int j = e / NN, rem = e - j * NN, r = rem / N, c = rem - r * N;
    lds[j * PL + r * NP + c] = load(j, r, c);
  }
  __syncthreads();
  if (tid < cnt * N) {
    const int j = tid / N, r = tid - j * N;
    cpx<T>* row = lds + j * PL + r * NP;
    cpx<T> v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = row[k];
    FFT<T, N, SIGN>::run(v);
#pragma unroll
    for (int k = 0; k < N; ++k) row[k] = v[k];
  }
  __syncthreads();
  if (tid < cnt * N) {
    const int j = tid / N, c = tid - j * N;
    cpx<T>* col = lds + j * PL + c;
    cpx<T> v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = col[k * NP];
    FFT<T, N, SIGN>::run(v);
#pragma unroll
    for (int k = 0; k < N; ++k) col[k * NP] = v[k];
  }
  __syncthreads();
  for (int e = tid; e < cnt * NN; e += NT) {
    const int j = e / NN, rem = e - j * NN, r = rem / N, c = rem - r * N;
    store(j, r, c, lds[j * PL + r * NP + c]);
  }
  __syncthreads();
}

// 1D FFTs along x of the N*N pencils (y, z); pencil idx = y*N + z.
template <typename T, int N, int NT, int SIGN, class Load, class Epi>
__device__ __forceinline__ void pencil_pass(Load load, Epi epi) {
  for (int idx = threadIdx.x; idx < N * N; idx += NT) {
    cpx<T> v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = load(k, idx);
    FFT<T, N, SIGN>::run(v);
    epi(idx, v);
  }
  __syncthreads();
}

// column c of row r in an N x N tile: XOR swizzle for power-of-two N (both
// row and column accesses conflict free), (c + r) mod N otherwise (a
// permutation of the row as well)
template <int N>
__device__ __forceinline__ int nyq_swizzle(int c, int r) {
  if constexpr ((N & (N - 1)) == 0) {
    return c ^ r;
  } else {
    const int s = c + r;
    return s >= N ? s - N : s;
  }
}

// f(integral_constant<int, I>) for I in [B, E), unrolled at compile time
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// G[k] = (first ? 0 : G[k]) + (w b[k]) v[k] for k < N, C values at a time,
// every load of a chunk issued before its first store.  G and b are pencil
// accessors (G[k] a reference, b[k] a real or complex value).  Written as one
// loop, each store could overlap the next loads as far as the compiler can
// tell, so every load would wait for the one before it to return.
template <typename T, int N, int C, class GA, class BA>
__device__ __forceinline__ void pencil_accumulate(const GA& G, const BA& b, T w, const cpx<T> (&v)[N], bool first) {
  static_assert(N % C == 0, "whole chunks");
  using BT = std::decay_t<decltype(b[0])>;
#pragma unroll
  for (int k0 = 0; k0 < N; k0 += C) {
    BT bv[C];
    cpx<T> gv[C];
#pragma unroll
    for (int k = 0; k < C; ++k) bv[k] = b[k0 + k];
    if (first) {
#pragma unroll
      for (int k = 0; k < C; ++k) gv[k] = cpx<T>{T(0), T(0)};
    } else {
#pragma unroll
      for (int k = 0; k < C; ++k) gv[k] = G[k0 + k];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < C; ++k) G[k0 + k] = gv[k] + (w * bv[k]) * v[k0 + k];
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Consecutive workgroups of a launch go round-robin to the 8 XCDs of an
// MI355X, each XCD with its own L2.  xcd_slot() renumbers block b so that the
// blocks of one XCD hold consecutive slots (persistent grids then give the
// work items of one point to one XCD, and its spectrum stays in that L2).  It
// is a bijection on [0, gridDim.x) whatever the hardware placement: only
// locality depends on it, never the result.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_slot() {
  const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
  return (nb % kXcds == 0) ? (b % kXcds) * (nb / kXcds) + b / kXcds : b;
}

template <int N>
struct Shape {
  static constexpr int NT = (N >= 16) ? 128 : 64;
  static constexpr int PB = NT / N;  // planes per slab batch
  // the Nyquist pass transforms 3 planes in one batch and the phase tables
  // are filled by 3N threads
  static_assert(PB >= 3, "slab batch must hold the three Nyquist planes");
  static_assert(3 * N <= NT, "phase tables need 3N threads");
};

template <typename T>
struct PrepArgs {
  const double* f;  // [npts][n] contiguous
  cpx<T>* F;        // [npts][n] natural spectral order, scaled by 1/n
  cpx<T>* Ft;       // optional (1 + i) F transposed to [kx][kz][ky] (register-resident gain)
  cpx<T>* Z;        // scratch [nslots][n]
  int npts;
};

template <typename T, int N>
__global__ __launch_bounds__(Shape<N>::NT) void k_vhs_prep(PrepArgs<T> a) {
  constexpr int NT = Shape<N>::NT, PB = Shape<N>::PB, NN = N * N, n = N * N * N;
  __shared__ cpx<T> lds[PB * N * (N + 1)];
  cpx<T>* Z = a.Z + static_cast<size_t>(blockIdx.x) * n;
  const T rn = static_cast<T>(1.0 / static_cast<double>(n));
  for (int pt = blockIdx.x; pt < a.npts; pt += gridDim.x) {
    const double* f = a.f + static_cast<size_t>(pt) * n;
    cpx<T>* F = a.F + static_cast<size_t>(pt) * n;
    cpx<T>* Ft = a.Ft ? a.Ft + static_cast<size_t>(pt) * n : nullptr;
    pencil_pass<T, N, NT, -1>(
        [&](int x, int idx) { return cpx<T>{static_cast<T>(f[x * NN + idx]), T(0)}; },
        [&](int idx, cpx<T>* v) {
#pragma unroll
          for (int k = 0; k < N; ++k) Z[k * NN + idx] = v[k];
        });
    for (int kx0 = 0; kx0 < N; kx0 += PB) {
      const int cnt = min(PB, N - kx0);
      slab_batch<T, N, NT, -1>(
          lds, cnt, [&](int j, int r, int c) { return Z[(kx0 + j) * NN + r * N + c]; },
          [&](int j, int r, int c, cpx<T> v) {
            F[(kx0 + j) * NN + r * N + c] = rn * v;
            if (Ft) Ft[(kx0 + j) * NN + c * N + r] = rn * cpx<T>{v.x - v.y, v.x + v.y};  // (1 + i) F
          });
    }
  }
}

// AoSoA <-> contiguous [pt][r] with pt = (elem - e0) * nupts + upt; f and out
// are the register base pointers (launchers defined next to the kernels).
int launch_aosoa_gather(const double* f, double* out, int nupts, int ldim, int vsize, int e0, int ne,
                        hipStream_t st);
int launch_aosoa_scatter(const double* in, double* Q, int nupts, int ldim, int vsize, int e0, int ne,
                         hipStream_t st);

// FP64 minimum-norm projection of Q[npts][N^3] onto the five collision
// invariants {1, c, |c|^2} (grid [-L, L]^3, weight cw); Ainv25 from
// invariant_gram_inverse.
int invariant_gram_inverse(int N, double L, double cw, double* Ainv25);
int launch_invariant_projection(double* Q, const double* Ainv25, double L, double cw, int N, int npts,
                                hipStream_t st);

// Per-kernel timing: event pairs recorded around every launch on the launch
// stream; resolved (and synchronised) only when queried.  H provides
// `timing`, `pool`, `pending`, `acc_ms[]` and `acc_n[]`.
template <class H>
hipEvent_t take_event(H* h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// record a start event for kernel class k (0 prep, 1 gain, 2 finalize, 3 project,
// 4 Nyquist planes)
template <class H>
struct KTimerT {
  H* h;
  int k;
  hipStream_t st;
  hipEvent_t a = nullptr;
  KTimerT(H* h_, int k_, hipStream_t st_) : h(h_), k(k_), st(st_) {
    if (h->timing) {
      a = take_event(h);
      if (a) (void)hipEventRecord(a, st);
    }
  }
  void stop() {
    if (!a) return;
    hipEvent_t b = take_event(h);
    if (!b) return;
    (void)hipEventRecord(b, st);
    h->pending.push_back({k, {a, b}});
    a = nullptr;
  }
};

template <class H>
void resolve_timing(H* h) {
  for (auto& pr : h->pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(pr.second.second);
    if (hipEventElapsedTime(&ms, pr.second.first, pr.second.second) == hipSuccess) {
      h->acc_ms[pr.first] += ms;
      h->acc_n[pr.first] += 1;
    }
    h->pool.push_back(pr.second.first);
    h->pool.push_back(pr.second.second);
  }
  h->pending.clear();
}

}  // namespace dgfs
