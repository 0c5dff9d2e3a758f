// Fast-spectral VHS collision operator Q(f, f) for gfx950.
//
// Algorithm (identical in exact arithmetic to the reference dataflow of
// frfs/solvers/dgfs/scattering.py:227-286 / vhs-gll.mako:103-266):
//
//   F      = FFT(f) / n                                   (prep kernel)
//   for every radial node p and direction q:
//     E(k)  = exp(i a_pq(k)),  a_pq = (pi/L)(qz_p/2)(l . sigma_q), separable
//             as ex(kx) ey(ky) ez(kz)
//     Y     = IFFT(F (E + i conj E))  = P+ + i P-, one complex transform for
//             the two real-part fields P+- = IFFT(exp(+-i a) F)
//     W     = IFFT(B), B the anti-Hermitian part of the two spectra, which
//             is supported on the three Nyquist planes only (the l = -N/2
//             row has no +N/2 partner, scattering.py:79); W carries the exact
//             imaginary parts I+- of P+- and costs three 2D transforms
//     S_p  += Re(P+ P-) = (Re Y + Im W)(Im Y - Re W) - Re W Im W
//   G^     = sum_p (2 qw_p sw / n) b1_p FFT(S_p)           (gain kernel)
//   x      = IFFT(G^ + i b2 F),  Q = prefac (Re x - f Im x) (finalize kernel)
//
// The reference computes t3 = IFFT(cos a F), t4 = IFFT(sin a F) and
// S = t3^2 + t4^2 (complex squares, vhs-gll.mako:160-164); t3^2 + t4^2 =
// (t3 + i t4)(t3 - i t4) = P+ P-.  Only Re S reaches Re G because b1 and the
// FFT of a real field are Hermitian, so the imaginary part of S is dropped.
// The sum over q is taken before the forward FFT (linearity), which turns the
// reference's 3K+3 transforms per point into K + Nrho + 3.
//
// Data movement: one workgroup owns a (point, radial group) work item and
// all of its 3D transforms.  A 3D transform is a slab pass (each kx-plane is
// 2D-transformed in LDS) plus a pencil pass (x-pencils in registers, one per
// thread, coalesced over z); the intermediate lives in a per-workgroup scratch
// region that the persistent grid reuses item after item, so it stays in the
// L2 / Infinity Cache instead of streaming to HBM.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "cplx.hpp"
#include "spectral.hpp"

namespace dgfs {

std::string& last_error_ref() {
  static thread_local std::string s;
  return s;
}


// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

template <typename T>
struct GainArgs {
  const cpx<T>* F;      // [npts][n]
  cpx<T>* Ghat;         // [npts*ngroups][n]
  cpx<T>* Z;            // scratch [nslots][n]
  T* S;                 // scratch [nslots][n]
  cpx<T>* W;            // scratch [nslots][3][N][N]
  const T* b1;          // [Nrho][n]
  const double* alpha;  // [Nrho] (pi/L) qz_p / 2
  const double* wgt;    // [Nrho] 2 qw_p sw / n
  const double* sz;     // [M][3]
  int nitems, ngroups, ppg, Nrho, M;
};

// One (point, radial group) work item of the multi-pass gain: LDS and phase
// tables come from the calling kernel, Z / S / W are this workgroup's scratch.
template <typename T, int N>
__device__ __forceinline__ void vhs_gain_item(const GainArgs<T>& a, int item, cpx<T>* lds, cpx<T> (*eph)[N]) {
  constexpr int NT = Shape<N>::NT, PB = Shape<N>::PB, NN = N * N, n = N * N * N, h = N / 2;
  const int tid = threadIdx.x;
  cpx<T>* Z = a.Z + static_cast<size_t>(blockIdx.x) * n;
  T* S = a.S + static_cast<size_t>(blockIdx.x) * n;
  cpx<T>* W = a.W + static_cast<size_t>(blockIdx.x) * 3 * NN;
  {
    const int pt = item / a.ngroups, g = item - pt * a.ngroups;
    const int p0 = g * a.ppg, p1 = min(p0 + a.ppg, a.Nrho);
    const cpx<T>* F = a.F + static_cast<size_t>(pt) * n;
    cpx<T>* G = a.Ghat + static_cast<size_t>(item) * n;
    for (int i = tid; i < n; i += NT) G[i] = cpx<T>{T(0), T(0)};
    for (int p = p0; p < p1; ++p) {
      for (int i = tid; i < n; i += NT) S[i] = T(0);
      const double alpha = a.alpha[p];
      for (int q = 0; q < a.M; ++q) {
        __syncthreads();
        if (tid < 3 * N) {
          const int d = tid / N, k = tid - d * N;
          const double l = (k < h) ? static_cast<double>(k) : static_cast<double>(k - N);
          double s, c;
          sincos(alpha * a.sz[q * 3 + d] * l, &s, &c);
          eph[d][k] = cpx<T>{static_cast<T>(c), static_cast<T>(s)};
        }
        __syncthreads();

        // Nyquist planes: B = (F/2)(conj D - i D), D = E - E~, where E~ takes
        // the conjugate factor on every axis whose index is the Nyquist row.
        slab_batch<T, N, NT, +1>(
            lds, 3,
            [&](int j, int r, int c) -> cpx<T> {
              int kx, ky, kz;
              if (j == 0) {
                kx = h; ky = r; kz = c;
              } else if (j == 1) {
                kx = r; ky = h; kz = c;
                if (kx == h) return cpx<T>{T(0), T(0)};
              } else {
                kx = r; ky = c; kz = h;
                if (kx == h || ky == h) return cpx<T>{T(0), T(0)};
              }
              const cpx<T> ex = eph[0][kx], ey = eph[1][ky], ez = eph[2][kz];
              const cpx<T> E = ex * ey * ez;
              const cpx<T> Et = (kx == h ? conj(ex) : ex) * (ky == h ? conj(ey) : ey) *
                                (kz == h ? conj(ez) : ez);
              const cpx<T> D = E - Et;
              const cpx<T> t = conj(D) - mul_i(D);
              return T(0.5) * (F[kx * NN + ky * N + kz] * t);
            },
            [&](int j, int r, int c, cpx<T> v) { W[j * NN + r * N + c] = v; });

        // main transform, yz-planes then x-pencils
        for (int kx0 = 0; kx0 < N; kx0 += PB) {
          const int cnt = min(PB, N - kx0);
          slab_batch<T, N, NT, +1>(
              lds, cnt,
              [&](int j, int ky, int kz) -> cpx<T> {
                const int kx = kx0 + j;
                const cpx<T> E = eph[0][kx] * eph[1][ky] * eph[2][kz];
                return F[kx * NN + ky * N + kz] * (E + mul_i(conj(E)));
              },
              [&](int j, int y, int z, cpx<T> v) { Z[(kx0 + j) * NN + y * N + z] = v; });
        }
        pencil_pass<T, N, NT, +1>(
            [&](int kx, int idx) { return Z[kx * NN + idx]; },
            [&](int idx, cpx<T>* v) {
              const int y = idx / N, z = idx - y * N;
              const cpx<T> w1 = W[idx];
              const T sy = (y & 1) ? T(-1) : T(1), szs = (z & 1) ? T(-1) : T(1);
#pragma unroll
              for (int x = 0; x < N; ++x) {
                const T sx = (x & 1) ? T(-1) : T(1);
                const cpx<T> w2 = W[NN + x * N + z], w3 = W[2 * NN + x * N + y];
                const T wr = sx * w1.x + sy * w2.x + szs * w3.x;
                const T wi = sx * w1.y + sy * w2.y + szs * w3.y;
                S[x * NN + idx] += (v[x].x + wi) * (v[x].y - wr) - wr * wi;
              }
            });
      }
      // forward FFT of S_p, weighted by 2 qw_p sw / n * b1_p, into G^
      pencil_pass<T, N, NT, -1>(
          [&](int x, int idx) { return cpx<T>{S[x * NN + idx], T(0)}; },
          [&](int idx, cpx<T>* v) {
#pragma unroll
            for (int k = 0; k < N; ++k) Z[k * NN + idx] = v[k];
          });
      const T wp = static_cast<T>(a.wgt[p]);
      const T* b1p = a.b1 + static_cast<size_t>(p) * n;
      for (int kx0 = 0; kx0 < N; kx0 += PB) {
        const int cnt = min(PB, N - kx0);
        slab_batch<T, N, NT, -1>(
            lds, cnt, [&](int j, int r, int c) { return Z[(kx0 + j) * NN + r * N + c]; },
            [&](int j, int r, int c, cpx<T> v) {
              const int o = (kx0 + j) * NN + r * N + c;
              G[o] = G[o] + (wp * b1p[o]) * v;
            });
      }
    }
  }
}

template <typename T, int N>
__global__ __launch_bounds__(Shape<N>::NT) void k_vhs_gain(GainArgs<T> a) {
  __shared__ cpx<T> lds[Shape<N>::PB * N * (N + 1)];
  __shared__ cpx<T> eph[3][N];
  for (int item = blockIdx.x; item < a.nitems; item += gridDim.x) vhs_gain_item<T, N>(a, item, lds, eph);
}

// Two velocity grids in one persistent launch: the work items of grid NA
// come first, then those of grid NB; each workgroup runs either kind with the
// scratch of the grid's own handle.  Both grids use Shape<N>::NT threads.
template <typename T, int NA, int NB>
__global__ __launch_bounds__(Shape<NA>::NT) void k_vhs_gain_pair(GainArgs<T> a, GainArgs<T> b) {
  static_assert(Shape<NA>::NT == Shape<NB>::NT, "paired grids need the same workgroup size");
  constexpr int LA = Shape<NA>::PB * NA * (NA + 1), LB = Shape<NB>::PB * NB * (NB + 1);
  __shared__ cpx<T> lds[LA > LB ? LA : LB];
  __shared__ cpx<T> eph[3][NA > NB ? NA : NB];
  auto* ea = reinterpret_cast<cpx<T>(*)[NA]>(&eph[0][0]);
  auto* eb = reinterpret_cast<cpx<T>(*)[NB]>(&eph[0][0]);
  const int total = a.nitems + b.nitems;
  for (int it = blockIdx.x; it < total; it += gridDim.x) {
    if (it < a.nitems)
      vhs_gain_item<T, NA>(a, it, lds, ea);
    else
      vhs_gain_item<T, NB>(b, it - a.nitems, lds, eb);
  }
}

// ---------------------------------------------------------------------------
// Nyquist-plane corrections for every (point, p, q): W = IFFT(B) on the three
// planes kx = N/2, ky = N/2 (kx != N/2), kz = N/2 (kx, ky != N/2), written to
// Wbuf[pt][p*M + q][3][N][N] (plane j indexed [r][c]: W1[y][z], W2[x][z],
// W3[x][y]).  Consumed by the register-resident gain kernel.
//
// With S the set of axes whose index is N/2 at a spectral point, E_S the
// product of their phase factors and e = a + ib the product of the others,
// the anti-Hermitian part 0.5 F (conj D - i D), D = E - E~, reduces to
//   B = (1 - i) F Im(E_S) (a - b),
// a real weight times F.  One wave transforms 64/N planes, N lanes each:
// lane c loads column c of its plane (coalesced; the kz = N/2 plane comes
// from the transposed spectrum), transforms along the rows, and the plane
// goes through a swizzled LDS tile (element (r, c) at r*N + (c ^ r), conflict
// free both ways) for the column transform and back for coalesced stores.
// ---------------------------------------------------------------------------
template <typename T>
struct NyqArgs {
  const cpx<T>* F;      // [npts][n]
  const cpx<T>* Ft;     // [npts][n] (1 + i) F transposed to [kx][kz][ky]
  cpx<T>* W;            // [npts][K][3][N][N]
  const cpx<T>* ephtab; // [K][3][N] phase factors
  long nplanes;         // npts * K * 3
  int Nrho, M;
};

template <int N>
struct NyqShape {
  static constexpr int PW = 64 / N;        // planes per wave (lanes past PW * N idle)
  static constexpr int PSTR = N * N + 16;  // LDS plane stride (planes on distinct banks)
};


template <typename T, int N>
__global__ __launch_bounds__(64) void k_vhs_nyq(NyqArgs<T> a) {
  auto sw = [](int c, int r) { return nyq_swizzle<N>(c, r); };
  constexpr int PW = NyqShape<N>::PW, PSTR = NyqShape<N>::PSTR, NN = N * N, n = NN * N, h = N / 2;
  static_assert(PW >= 1 && PW * N <= 64, "one wave per PW planes");
  __shared__ cpx<T> buf[PW * PSTR];
  __shared__ cpx<T> erow[PW][N];  // row-axis factor, 1 at index N/2
  __shared__ cpx<T> enyq[PW][2];  // row- and column-axis factors at index N/2
  // lanes past PW * N (N not dividing 64) mirror the last plane's lanes
  // inactively: they take part in the barriers only
  const int lane = threadIdx.x, wl = lane / N;
  const bool spare = wl >= PW;
  const int w = spare ? PW - 1 : wl, c = spare ? lane - (PW - 1) * N - N : lane - wl * N;
  const int K = a.Nrho * a.M;
  cpx<T>* P = buf + w * PSTR;
  for (long base = static_cast<long>(blockIdx.x) * PW; base < a.nplanes;
       base += static_cast<long>(gridDim.x) * PW) {
    const long pl = base + w;
    const bool active = !spare && pl < a.nplanes;
    const long item = active ? pl / 3 : 0;
    const int j = active ? static_cast<int>(pl - item * 3) : 0;
    const int pt = static_cast<int>(item / K), pq = static_cast<int>(item - static_cast<long>(pt) * K);
    const int ra = (j == 0) ? 1 : 0;  // row axis
    const int cb = (j == 2) ? 1 : 2;  // column axis
    const cpx<T>* E = a.ephtab + static_cast<size_t>(pq) * 3 * N;
    const cpx<T> er = E[ra * N + c], ec = E[cb * N + c], en = E[j * N + h];
    if (!spare) {
      erow[w][c] = (c == h) ? cpx<T>{T(1), T(0)} : er;
      if (c == h) {
        enyq[w][0] = er;
        enyq[w][1] = ec;
      }
    }
    __syncthreads();
    // Im(E_S) for the four Nyquist patterns (row N/2?, column N/2?)
    const cpx<T> ya = enyq[w][0], yb = enyq[w][1];
    const T s00 = en.y, s10 = (en * ya).y, s01 = (en * yb).y, s11 = (en * ya * yb).y;
    const bool ch = (c == h);
    T sc0 = ch ? s01 : s00;  // rows != N/2
    T sc1 = ch ? s11 : s10;  // row N/2
    if (j >= 1) sc1 = T(0);            // row kx = N/2 belongs to plane 0
    if (j == 2 && ch) sc0 = T(0);      // column ky = N/2 belongs to plane 1
    // column factor b' (1 at N/2): m(r, c) = Re(a' b') - Im(a' b') = ar u - ai v
    const T br = ch ? T(1) : ec.x, bi = ch ? T(0) : ec.y;
    const T u = br - bi, v = bi + br;
    const cpx<T>* src = (j == 2 ? a.Ft : a.F) + static_cast<size_t>(pt) * n + (j == 0 ? h * NN : h * N) + c;
    const int stride = (j == 0) ? N : NN;
    cpx<T> x[N];
#pragma unroll
    for (int r = 0; r < N; ++r) {
      const cpx<T> e = erow[w][r];
      const T m = (r == h ? sc1 : sc0) * (e.x * u - e.y * v);
      cpx<T> fv = active ? src[r * stride] : cpx<T>{T(0), T(0)};
      if (j == 2) fv = cpx<T>{T(0.5) * (fv.x + fv.y), T(0.5) * (fv.y - fv.x)};  // undo the (1 + i)
      x[r] = m * fv;
    }
    FFT<T, N, +1>::run(x);  // along the rows
    if (!spare) {
#pragma unroll
      for (int r = 0; r < N; ++r) P[r * N + sw(c, r)] = x[r];
    }
    __syncthreads();
    // lane c now owns row c of the plane
    if (!spare) {
#pragma unroll
      for (int k = 0; k < N; ++k) x[k] = P[c * N + sw(k, c)];
    }
    FFT<T, N, +1>::run(x);  // along the columns
    if (!spare) {
#pragma unroll
      for (int k = 0; k < N; ++k) P[c * N + sw(k, c)] = x[k];
    }
    __syncthreads();
    if (active) {
      cpx<T>* Wp = a.W + static_cast<size_t>(item) * 3 * NN + j * NN + c;
#pragma unroll
      for (int r = 0; r < N; ++r) {
        const cpx<T> y = P[r * N + sw(c, r)];
        // times (1 - i); non-temporal stores, as the planes are read back
        // only by the gain kernel after the whole launch (4.3 -> 4.0 ms per
        // 4096 points at 32^3 FP32 on MI355X)
        using V2 = T __attribute__((ext_vector_type(2)));
        const V2 wv = {y.x + y.y, y.y - y.x};
        __builtin_nontemporal_store(wv, reinterpret_cast<V2*>(Wp + r * N));
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Register-resident gain kernel.  A workgroup of N*N/2 threads holds one whole
// N^3 complex field in registers, two pencils per thread; thread t belongs
// to plane group g = t / N (N threads, one kx plane per slot) and lane
// l = t % N.  Per (p, q):
//   load   F plane kx = g + (N/2) j (slot j) through LDS, z-pencil ky = l
//   z-IFFT, in-plane LDS transpose, y-IFFT       -> y-pencil (kx, z = l)
//   two LDS rounds (kx halves) exchange kx <-> y -> x-pencil (y = g + (N/2) j, z = l)
//   x-IFFT, Re(P+ P-) with the precomputed Nyquist planes, accumulate S_p
// S_p stays in registers across q; at the end of p it is forward transformed
// by the reverse sequence and accumulated into G^ with weight w_p b1_p.
// LDS planes are swizzled (column c stored at c ^ r) so row and column
// accesses are both bank-conflict free.
// ---------------------------------------------------------------------------
template <typename T>
struct GainRRArgs {
  const cpx<T>* Ft;     // [npts][n] (1 + i) F transposed to [kx][kz][ky]
  const cpx<T>* W;      // [npts][K][3][N][N]
  cpx<T>* Ghat;         // [npts*ngroups][n]
  const T* b1;          // [Nrho][n]
  const double* alpha;  // [Nrho]
  const double* wgt;    // [Nrho]
  const double* sz;     // [M][3]
  const cpx<T>* ephtab; // [K][3][N] phase factors
  int nitems, ngroups, ppg, Nrho, M;
};

template <int N>
struct RRShape {
  static constexpr int NT = N * N / 2;
  static constexpr int G = N / 2;  // plane groups
};

// stage helpers of the register-resident kernel (force-inlined; scalars by
// value, register arrays by reference so they stay in VGPRs).  LDS layout:
//   Xb : N/4 padded N x (N+1) planes -- the in-plane transposes of half the
//        plane groups at a time, and the N/2 x N x N/2... exchange buffer
//   S0 : the slot-0 accumulators, S0[x][thread]
//   Wl : the three Nyquist planes of the current (p, q)
// Every LDS access is a per-thread base plus a compile-time offset; padded
// rows make row and column accesses bank-conflict free.
template <int N>
struct RRLds {
  static constexpr int plane = N * (N + 1);           // complex
  static constexpr int xbuf = (N / 4) * plane;        // complex; >= (N/2)*N*(N/2)/... checked below
  static constexpr int ex = (N / 2) * N * (N / 2);    // exchange sub-round: G planes x N z x G y... complex
  static constexpr int xsize = (xbuf > ex) ? xbuf : ex;
  static constexpr int wl = 3 * N * N;                // complex
  static constexpr int s0 = N * N * N / 2;            // real (one per thread and x)
};

// LDS writes and reads of one wave complete before it goes on: the in-plane
// transposes of the wave-local form exchange values between the lanes of
// one wave only, so no workgroup barrier is needed, and the wave waits for
// its own LDS operations only
__device__ __forceinline__ void lds_wave_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Copy NCH 16-byte chunks from global memory to LDS with the load unit
// (global_load_lds_dwordx4: lane i of a wave writes the wave's LDS base +
// 16 i, so chunk c goes to dst + 16 c), without passing through registers.
// The copy is in flight until the issuing wave waits for its vector memory
// counter: call lds_copy_wait() and then a workgroup barrier before any wave
// reads dst.  src and dst 16-byte aligned; NT a multiple of 64.
template <int NT, int NCH>
__device__ __forceinline__ void lds_copy16(const void* src, void* dst) {
  static_assert(NT % 64 == 0, "whole waves");
  const int tid = threadIdx.x, wbase = tid & ~63;
#pragma unroll
  for (int k = 0; k < (NCH + NT - 1) / NT; ++k) {
    const int c = tid + k * NT;
    if (NCH % NT == 0 || c < NCH)
      __builtin_amdgcn_global_load_lds(static_cast<const char*>(src) + 16 * static_cast<size_t>(c),
                                       (void __attribute__((address_space(3)))*)(static_cast<char*>(dst) +
                                                                                 16 * (wbase + k * NT)),
                                       16, 0, 0);
  }
}
__device__ __forceinline__ void lds_copy_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// In-plane transpose of the pencils of plane group g (lane l of the group)
// through LDS.  Workgroup form (WL false): plane groups [half*G/2,
// (half+1)*G/2) take turns on G/2 padded planes between workgroup barriers.
// Wave-local form (WL true): the 64/N plane groups of a wave take turns on
// the wave's own padded plane (NT/64 planes in all), with wave-level waits
// only.  ROWS: the thread writes its pencil as row l and reads column l
// (z -> y); else it writes column l and reads row l (y -> z).
template <typename T, int N, int G, bool WL, bool ROWS>
__device__ __forceinline__ void rr_plane_transpose(cpx<T> (&v)[N], cpx<T>* Xb, int g, int l) {
  constexpr int NP = N + 1;
  if constexpr (WL) {
    constexpr int GPW = 64 / N;
    static_assert(64 % N == 0, "plane groups must not straddle waves");
    const int w = g / GPW, sub = g - (g / GPW) * GPW;
    cpx<T>* Pc = Xb + w * (N * NP) + l;
    cpx<T>* Pr = Xb + w * (N * NP) + l * NP;
#pragma unroll
    for (int r = 0; r < GPW; ++r) {
      if (sub == r) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if constexpr (ROWS) Pr[k] = v[k]; else Pc[k * NP] = v[k];
        }
      }
      lds_wave_sync();
      if (sub == r) {
#pragma unroll
        for (int k0 = 0; k0 < N; k0 += 8) {
#pragma unroll
          for (int k = k0; k < k0 + 8; ++k) v[k] = ROWS ? Pc[k * NP] : Pr[k];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      lds_wave_sync();
    }
  } else {
    constexpr int GH = G / 2;
    const int slot = g % GH;
    cpx<T>* Pc = Xb + slot * (N * NP) + l;
    cpx<T>* Pr = Xb + slot * (N * NP) + l * NP;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const bool mine = (g / GH) == half;
      __syncthreads();
      if (mine) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if constexpr (ROWS) Pr[k] = v[k]; else Pc[k * NP] = v[k];
        }
      }
      __syncthreads();
      if (mine) {
#pragma unroll
        for (int k0 = 0; k0 < N; k0 += 8) {
#pragma unroll
          for (int k = k0; k < k0 + 8; ++k) v[k] = ROWS ? Pc[k * NP] : Pr[k];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  }
}

// inverse z then y transform of plane kx for the thread (group g, lane l)
template <typename T, int N, int G, bool WL>
__device__ __forceinline__ void rr_zy_inverse(cpx<T> (&v)[N], const cpx<T>* __restrict__ Ft, cpx<T>* Xb,
                                              const cpx<T> (*eph)[N], int kx, int g, int l) {
  // eph[0..2]: phase factors of the x, y, z axes; eph[3][k] = (Re + Im, Re - Im) of eph[2][k]
  constexpr int NN = N * N;
  // z-pencil (kx, ky = l) from the transposed spectrum Ft[kx][kz][ky].  The
  // loads of a pencil use at most two base addresses and immediate offsets
  // (4 KB reach): the second base is made opaque so the compiler does not
  // precompute one 64-bit address per kz outside the (p, q) loop.
  constexpr int KB = (4096 / static_cast<int>(N * sizeof(cpx<T>))) < N ? (4096 / static_cast<int>(N * sizeof(cpx<T>))) : N;
  const cpx<T>* P0 = Ft + kx * NN + l;
  int second = KB * N;
  asm volatile("" : "+v"(second));
  const cpx<T>* P1 = P0 + second;
#pragma unroll
  for (int kz = 0; kz < N; ++kz) v[kz] = (kz < KB) ? P0[kz * N] : P1[(kz - KB) * N];
  // F (E + i conj E) = (Re E + Im E) (1 + i) F with E = exy ez(kz); the
  // spectrum already carries the (1 + i), and Re E + Im E = A.x (B.x + B.y)
  // + A.y (B.x - B.y) with A = exy, B = ez, the pair taken from eph[3]
  const cpx<T> exy = eph[0][kx] * eph[1][l];
#pragma unroll
  for (int k0 = 0; k0 < N; k0 += 8) {
#pragma unroll
    for (int kz = k0; kz < k0 + 8; ++kz) {
      const cpx<T> pm = eph[3][kz];
      const T s = exy.x * pm.x + exy.y * pm.y;
      v[kz] = s * v[kz];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  FFT<T, N, +1, true>::run(v);
  __builtin_amdgcn_sched_barrier(0);
  rr_plane_transpose<T, N, G, WL, true>(v, Xb, g, l);
  FFT<T, N, +1, true>::run(v);
  __builtin_amdgcn_sched_barrier(0);
}

// x exchange for target slot r (y = g + G r), in two sub-rounds: sub-round sr
// moves the y-half r of plane slot sr (kx = g + G sr) through the buffer
// [kxl][yy][z]; the caller places the barriers, so that a write of the next
// target's first sub-round can overlap the previous target's x-IFFT
template <typename T, int N, int G>
__device__ __forceinline__ void rr_x_write(const cpx<T> (&v)[N], cpx<T>* Xb, int r, int g, int l) {
  cpx<T>* Xw = Xb + (g * G) * N + l;  // [kxl][yy][z]
#pragma unroll
  for (int yy = 0; yy < G; ++yy) Xw[yy * N] = v[G * r + yy];
}

template <typename T, int N, int G>
__device__ __forceinline__ void rr_x_read(cpx<T> (&u)[N], const cpx<T>* Xb, int sr, int g, int l) {
  const cpx<T>* Xr = Xb + g * N + l;  // read (kxl, yy = g, z = l)
#pragma unroll
  for (int k0 = 0; k0 < G; k0 += 8) {
#pragma unroll
    for (int kk = k0; kk < k0 + 8; ++kk) u[G * sr + kk] = Xr[kk * G * N];
    __builtin_amdgcn_sched_barrier(0);
  }
}

// x-IFFT of the gathered pencil (y = g + G r, z = l) and Re(P+ P-) with the
// Nyquist-plane corrections, returned in val[x]
template <typename T, int N, int G>
__device__ __forceinline__ void rr_x_compute(T (&val)[N], cpx<T> (&u)[N], const cpx<T>* Wl, int r, int g, int l) {
  constexpr int NN = N * N;
  FFT<T, N, +1, true>::run(u);
  __builtin_amdgcn_sched_barrier(0);
  const int y = g + G * r, z = l;
  const cpx<T> w1 = Wl[y * N + z];
  const cpx<T>* W2 = Wl + NN + z;
  const cpx<T>* W3 = Wl + 2 * NN + y;
  const T sy = (y & 1) ? T(-1) : T(1), szs = (z & 1) ? T(-1) : T(1);
#pragma unroll
  for (int x0 = 0; x0 < N; x0 += 8) {
#pragma unroll
    for (int x = x0; x < x0 + 8; ++x) {
      const T sx = (x & 1) ? T(-1) : T(1);
      const cpx<T> w2 = W2[x * N], w3 = W3[x * N];
      const T wr = sx * w1.x + sy * w2.x + szs * w3.x;
      const T wi = sx * w1.y + sy * w2.y + szs * w3.y;
      val[x] = (u[x].x + wi) * (u[x].y - wr) - wr * wi;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// x-FFT of the slot-r accumulator and the reverse exchange into the y-half r
// of both plane slots (two sub-rounds); TF: tangent-form butterflies
template <typename T, int N, int G, bool TF = true>
__device__ __forceinline__ void rr_x_forward(const T (&sacc)[N], cpx<T> (&v0)[N], cpx<T> (&v1)[N], cpx<T>* Xb,
                                             int r, int g, int l) {
  cpx<T> u[N];
#pragma unroll
  for (int x = 0; x < N; ++x) u[x] = cpx<T>{sacc[x], T(0)};
  FFT<T, N, -1, TF>::run(u);
  __builtin_amdgcn_sched_barrier(0);
  cpx<T>* Xw = Xb + g * N + l;         // write (kxl, yy = g, z = l)
  const cpx<T>* Xr = Xb + (g * G) * N + l;
#pragma unroll
  for (int sr = 0; sr < 2; ++sr) {
    cpx<T> (&v)[N] = sr ? v1 : v0;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < G; ++kk) Xw[kk * G * N] = u[G * sr + kk];
    __syncthreads();
#pragma unroll
    for (int y0 = 0; y0 < G; y0 += 8) {
#pragma unroll
      for (int yy = y0; yy < y0 + 8; ++yy) v[G * r + yy] = Xr[yy * N];
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// forward y then z transform of plane kx, weighted accumulation into G^;
// sync: a workgroup barrier first (the wave-local transpose must not
// overwrite the exchange buffer while other waves still read it)
template <typename T, int N, int G, bool WL>
__device__ __forceinline__ void rr_yz_forward(cpx<T> (&v)[N], cpx<T>* Xb, int kx, int g, int l, T wp,
                                              const T* __restrict__ b1p, cpx<T>* Gh, bool first, bool sync) {
  constexpr int NN = N * N;
  FFT<T, N, -1, true>::run(v);
  __builtin_amdgcn_sched_barrier(0);
  if (WL && sync) __syncthreads();
  rr_plane_transpose<T, N, G, WL, false>(v, Xb, g, l);
  FFT<T, N, -1, true>::run(v);
  __builtin_amdgcn_sched_barrier(0);
  const size_t o = static_cast<size_t>(kx) * NN + l * N;
  const T* b = b1p + o;
  cpx<T>* gp = Gh + o;
#pragma unroll
  for (int k0 = 0; k0 < N; k0 += 8) {
    if (first) {
#pragma unroll
      for (int kz = k0; kz < k0 + 8; ++kz) gp[kz] = (wp * b[kz]) * v[kz];
    } else {
#pragma unroll
      for (int kz = k0; kz < k0 + 8; ++kz) gp[kz] = gp[kz] + (wp * b[kz]) * v[kz];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// INL: the three Nyquist-plane transforms of every (p, q) are computed in the
// kernel (from the transposed spectrum, into LDS) instead of being read from
// the buffer k_vhs_nyq writes
// WL: in-plane transposes wave-local (rr_plane_transpose)
template <typename T, int N, bool INL, bool WL = false>
__global__ __launch_bounds__(RRShape<N>::NT) void k_vhs_gain_rr(GainRRArgs<T> a) {
  constexpr int NT = RRShape<N>::NT, G = RRShape<N>::G, NN = N * N, n = N * N * N, h = N / 2;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* Xb = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* Wl = Xb + RRLds<N>::xsize;
  T* S0 = reinterpret_cast<T*>(Wl + RRLds<N>::wl) + threadIdx.x;  // S0[x * NT]
  __shared__ cpx<T> eph[4][N];  // x, y, z phase factors and the z pair (Re + Im, Re - Im)
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  const int K = a.Nrho * a.M;

  cpx<T> v0[N], v1[N];
  T s1[N];

  for (int item = xcd_slot(); item < a.nitems; item += gridDim.x) {
    const int pt = item / a.ngroups, grp = item - pt * a.ngroups;
    const int p0 = grp * a.ppg, p1 = min(p0 + a.ppg, a.Nrho);
    const cpx<T>* Ft = a.Ft + static_cast<size_t>(pt) * n;
    cpx<T>* Gh = a.Ghat + static_cast<size_t>(item) * n;
    for (int p = p0; p < p1; ++p) {
#pragma unroll
      for (int x = 0; x < N; ++x) {
        s1[x] = T(0);
        S0[x * NT] = T(0);
      }
      for (int q = 0; q < a.M; ++q) {
        __syncthreads();
        if (tid < 3 * N) {
          const cpx<T> e = a.ephtab[(static_cast<size_t>(p) * a.M + q) * 3 * N + tid];
          (&eph[0][0])[tid] = e;
          if (tid >= 2 * N) eph[3][tid - 2 * N] = cpx<T>{e.x + e.y, e.x - e.y};
        }
        if constexpr (INL) {
          // B = (F/2)(conj D - i D) on the planes kx = h, ky = h (kx != h),
          // kz = h (kx, ky != h), D = E - E~ (as k_vhs_nyq), then 2D IFFTs
          __syncthreads();
          for (int i = tid; i < 3 * NN; i += NT) {
            const int j = i / NN, rem = i - j * NN, r = rem / N, c = rem - r * N;
            int kx, ky, kz;
            bool zero = false;
            if (j == 0) {
              kx = h; ky = r; kz = c;
            } else if (j == 1) {
              kx = r; ky = h; kz = c;
              zero = (kx == h);
            } else {
              kx = r; ky = c; kz = h;
              zero = (kx == h || ky == h);
            }
            cpx<T> v{T(0), T(0)};
            if (!zero) {
              const cpx<T> ex = eph[0][kx], ey = eph[1][ky], ez = eph[2][kz];
              const cpx<T> E = ex * ey * ez;
              const cpx<T> Et = (kx == h ? conj(ex) : ex) * (ky == h ? conj(ey) : ey) * (kz == h ? conj(ez) : ez);
              const cpx<T> D = E - Et;
              const cpx<T> fr = Ft[kx * NN + kz * N + ky];  // (1 + i) F
              const cpx<T> f0{T(0.5) * (fr.x + fr.y), T(0.5) * (fr.y - fr.x)};
              v = T(0.5) * (f0 * (conj(D) - mul_i(D)));
            }
            Wl[i] = v;
          }
          __syncthreads();
          if (tid < 3 * N) {  // along c
            cpx<T> w[N];
            cpx<T>* row = Wl + tid * N;
#pragma unroll
            for (int k = 0; k < N; ++k) w[k] = row[k];
            FFT<T, N, +1>::run(w);
#pragma unroll
            for (int k = 0; k < N; ++k) row[k] = w[k];
          }
          __syncthreads();
          if (tid < 3 * N) {  // along r
            const int j = tid / N, c = tid - j * N;
            cpx<T> w[N];
            cpx<T>* col = Wl + j * NN + c;
#pragma unroll
            for (int k = 0; k < N; ++k) w[k] = col[k * N];
            FFT<T, N, +1>::run(w);
#pragma unroll
            for (int k = 0; k < N; ++k) col[k * N] = w[k];
          }
          __syncthreads();
        } else {
          // this (p, q)'s Nyquist planes into LDS by the load unit (the
          // compiler completes the copy at the next barrier; 67.9 -> 66.7 ms
          // per 4096-point launch at 32^3 on MI355X against staging them
          // through registers and an LDS store)
          const cpx<T>* Wq = a.W + (static_cast<size_t>(pt) * K + p * a.M + q) * 3 * NN;
          lds_copy16<NT, 3 * NN * static_cast<int>(sizeof(cpx<T>)) / 16>(Wq, Wl);
          __syncthreads();  // eph is in place
        }
        rr_zy_inverse<T, N, G, WL>(v0, Ft, Xb, eph, g, g, l);
        rr_zy_inverse<T, N, G, WL>(v1, Ft, Xb, eph, g + G, g, l);
        T val[N];
        cpx<T> u[N];
        if constexpr (!INL) lds_copy_wait();  // the Nyquist planes have landed (read after the next barrier)
        // target slot 0 (y = g): both sub-rounds
        __syncthreads();
        rr_x_write<T, N, G>(v0, Xb, 0, g, l);
        __syncthreads();
        rr_x_read<T, N, G>(u, Xb, 0, g, l);
        __syncthreads();
        rr_x_write<T, N, G>(v1, Xb, 0, g, l);
        __syncthreads();
        rr_x_read<T, N, G>(u, Xb, 1, g, l);
        __syncthreads();
        // target slot 1's first sub-round is written while slot 0 transforms
        rr_x_write<T, N, G>(v0, Xb, 1, g, l);
        rr_x_compute<T, N, G>(val, u, Wl, 0, g, l);
#pragma unroll
        for (int x = 0; x < N; ++x) S0[x * NT] += val[x];
        __syncthreads();
        rr_x_read<T, N, G>(u, Xb, 0, g, l);
        __syncthreads();
        rr_x_write<T, N, G>(v1, Xb, 1, g, l);
        __syncthreads();
        rr_x_read<T, N, G>(u, Xb, 1, g, l);
        rr_x_compute<T, N, G>(val, u, Wl, 1, g, l);
#pragma unroll
        for (int x = 0; x < N; ++x) s1[x] += val[x];
      }
      {
        T s0r[N];
#pragma unroll
        for (int x = 0; x < N; ++x) s0r[x] = S0[x * NT];
        rr_x_forward<T, N, G>(s0r, v0, v1, Xb, 0, g, l);
      }
      rr_x_forward<T, N, G>(s1, v0, v1, Xb, 1, g, l);
      const T wp = static_cast<T>(a.wgt[p]);
      const T* b1p = a.b1 + static_cast<size_t>(p) * n;
      rr_yz_forward<T, N, G, WL>(v0, Xb, g, g, l, wp, b1p, Gh, p == p0, true);
      rr_yz_forward<T, N, G, WL>(v1, Xb, g + G, g, l, wp, b1p, Gh, p == p0, false);
    }
  }
}

// ---------------------------------------------------------------------------
// Register-resident gain kernel, one pencil per thread (N whose N^3 complex
// field fits in LDS: 16, 24).  A workgroup of N*N threads owns one work item;
// thread t = g N + l holds one pencil.  Per (p, q):
//   z-pencil (kx = g, ky = l) from the transposed (1 + i)-folded spectrum,
//   phase weight, z-IFFT; the kx plane goes through LDS ([ky][z], rows
//   padded to N+1) to give the y-pencil (kx = g, z = l), y-IFFT; the whole
//   field goes through LDS once ([kx][y][z]) to give the x-pencil (y = g,
//   z = l), x-IFFT, Re(P+ P-) with the Nyquist-plane corrections, S_p += in
//   registers.  At the end of p the reverse sequence forward transforms S_p
//   and accumulates w_p b1_p FFT(S_p) into G^.  Same algebra and tables as
//   k_vhs_gain_rr; every LDS access is a lane-consecutive row.
// ---------------------------------------------------------------------------
template <int N>
struct RR1Shape {
  static constexpr int NT = N * N;
  static constexpr int PP = N + 1;          // padded row pitch of the kx planes
  static constexpr int field = N * N * PP;  // complex; holds the N^3 exchange too
  static constexpr int wl = 3 * N * N;      // complex, Nyquist planes
  // kx stride of the field exchange: N (N + 1) complex, so that the rows of
  // consecutive kx that one wave touches fall on different LDS banks
  static constexpr int KS = N * PP;
  static_assert(field >= (N - 1) * KS + N * N, "exchange must fit the plane buffer");
};

template <typename T, int N>
__global__ __launch_bounds__(RR1Shape<N>::NT) void k_vhs_gain_rr1(GainRRArgs<T> a) {
  constexpr int NT = RR1Shape<N>::NT, PP = RR1Shape<N>::PP, NN = N * N, n = NN * N;
  constexpr int KS = RR1Shape<N>::KS;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* Wl = X + RR1Shape<N>::field;
  __shared__ cpx<T> eph[4][N];  // x, y, z phase factors and the z pair (Re + Im, Re - Im)
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  const int K = a.Nrho * a.M;
  cpx<T>* Xp = X + g * N * PP;  // kx plane g
  T sacc[N];
  cpx<T> v[N];
  for (int item = xcd_slot(); item < a.nitems; item += gridDim.x) {
    const int pt = item / a.ngroups, grp = item - pt * a.ngroups;
    const int p0 = grp * a.ppg, p1 = min(p0 + a.ppg, a.Nrho);
    const cpx<T>* Ft = a.Ft + static_cast<size_t>(pt) * n;
    cpx<T>* Gh = a.Ghat + static_cast<size_t>(item) * n;
    for (int p = p0; p < p1; ++p) {
#pragma unroll
      for (int x = 0; x < N; ++x) sacc[x] = T(0);
      for (int q = 0; q < a.M; ++q) {
        __syncthreads();  // readers of X, Wl, eph from the previous (p, q) are done
        if (tid < 3 * N) {
          const cpx<T> e = a.ephtab[(static_cast<size_t>(p) * a.M + q) * 3 * N + tid];
          (&eph[0][0])[tid] = e;
          if (tid >= 2 * N) eph[3][tid - 2 * N] = cpx<T>{e.x + e.y, e.x - e.y};
        }
        {
          // this (p, q)'s Nyquist planes through registers: their loads overlap
          // the z-pencil loads below (a copy by the load unit, which the
          // compiler completes at the next barrier, measured 2.5 % slower
          // here on MI355X)
          const cpx<T>* Wq = a.W + (static_cast<size_t>(pt) * K + p * a.M + q) * 3 * NN;
          static_assert((3 * NN) % NT == 0, "Nyquist planes split evenly over the workgroup");
          cpx<T> wv[3 * NN / NT];
#pragma unroll
          for (int k = 0; k < 3 * NN / NT; ++k) wv[k] = Wq[tid + k * NT];
#pragma unroll
          for (int k = 0; k < 3 * NN / NT; ++k) Wl[tid + k * NT] = wv[k];
        }
        // z-pencil (kx = g, ky = l): two base addresses with immediate offsets
        {
          constexpr int KB = (4096 / static_cast<int>(N * sizeof(cpx<T>))) < N
                                 ? (4096 / static_cast<int>(N * sizeof(cpx<T>))) : N;
          const cpx<T>* P0 = Ft + g * NN + l;
          int second = KB * N;
          asm volatile("" : "+v"(second));
          const cpx<T>* P1 = P0 + second;
#pragma unroll
          for (int kz = 0; kz < N; ++kz) v[kz] = (kz < KB) ? P0[kz * N] : P1[(kz - KB) * N];
        }
        __syncthreads();  // eph and Wl are in place
        {
          const cpx<T> exy = eph[0][g] * eph[1][l];
#pragma unroll
          for (int kz = 0; kz < N; ++kz) {
            const cpx<T> pm = eph[3][kz];
            v[kz] = (exy.x * pm.x + exy.y * pm.y) * v[kz];
          }
        }
        FFT<T, N, +1, true>::run(v);  // z
#pragma unroll
        for (int z = 0; z < N; ++z) Xp[l * PP + z] = v[z];
        __syncthreads();
#pragma unroll
        for (int ky = 0; ky < N; ++ky) v[ky] = Xp[ky * PP + l];
        FFT<T, N, +1, true>::run(v);  // y: (kx = g, y, z = l)
        __syncthreads();
#pragma unroll
        for (int y = 0; y < N; ++y) X[g * KS + y * N + l] = v[y];
        __syncthreads();
#pragma unroll
        for (int kx = 0; kx < N; ++kx) v[kx] = X[kx * KS + g * N + l];
        FFT<T, N, +1, true>::run(v);  // x: (x, y = g, z = l)
        {
          const cpx<T> w1 = Wl[g * N + l];
          const cpx<T>* W2 = Wl + NN + l;
          const cpx<T>* W3 = Wl + 2 * NN + g;
          const T sy = (g & 1) ? T(-1) : T(1), szs = (l & 1) ? T(-1) : T(1);
#pragma unroll
          for (int x = 0; x < N; ++x) {
            const T sx = (x & 1) ? T(-1) : T(1);
            const cpx<T> w2 = W2[x * N], w3 = W3[x * N];
            const T wr = sx * w1.x + sy * w2.x + szs * w3.x;
            const T wi = sx * w1.y + sy * w2.y + szs * w3.y;
            sacc[x] += (v[x].x + wi) * (v[x].y - wr) - wr * wi;
          }
        }
      }
      // forward transform of S_p and accumulation into G^
#pragma unroll
      for (int x = 0; x < N; ++x) v[x] = cpx<T>{sacc[x], T(0)};
      FFT<T, N, -1, true>::run(v);  // x: (kx, y = g, z = l)
      __syncthreads();
#pragma unroll
      for (int kx = 0; kx < N; ++kx) X[kx * KS + g * N + l] = v[kx];
      __syncthreads();
#pragma unroll
      for (int y = 0; y < N; ++y) v[y] = X[g * KS + y * N + l];
      FFT<T, N, -1, true>::run(v);  // y: (kx = g, ky, z = l)
      __syncthreads();
#pragma unroll
      for (int ky = 0; ky < N; ++ky) Xp[ky * PP + l] = v[ky];
      __syncthreads();
#pragma unroll
      for (int z = 0; z < N; ++z) v[z] = Xp[l * PP + z];
      FFT<T, N, -1, true>::run(v);  // z: (kx = g, ky = l, kz)
      const T wp = static_cast<T>(a.wgt[p]);
      const size_t o = static_cast<size_t>(g) * NN + l * N;
      const T* b = a.b1 + static_cast<size_t>(p) * n + o;
      pencil_accumulate<T, N, N>(Gh + o, b, wp, v, p == p0);
    }
  }
}

// ---------------------------------------------------------------------------
// Slab gain kernel: the FP64 form for fields that do not fit on one CU (a
// 32^3 FP64 complex field is 512 KB, the whole register file).  A workgroup
// owns a (point, radial group, slab) item; slab s holds the XS = N/NS planes
// x = s + NS j (j < XS) of the inverse transforms, so a slab of XS planes is
// all that sits in LDS.  Per (p, q):
//   x-pass  pencils (kz, ky) of the (1 + i)-folded spectrum times the real
//           phase factor Re E + Im E.  Only the outputs x = s + NS j of the
//           N-point inverse DFT are needed: with kx = r + XS m,
//             Y(s + NS j) = sum_r w_XS^{r j} w_N^{r s} sum_m w_NS^{m s} u(r + XS m),
//           a fold over m, a twiddle and an XS-point inverse FFT
//           (w_K = exp(2 pi i / K)); plane j, row ky, column kz of LDS.
//   z-pass  one row (plane j, ky) per thread, inverse FFT along kz.
//   y-pass  one column (plane j, z) per thread, inverse FFT along ky, then
//           Re(P+ P-) with the Nyquist-plane corrections into S_p[y], which
//           stays in the thread's registers over q.
// At the end of p the slab of S_p is forward transformed along y and z and
// stored as T[point][p][x][ky][kz]; k_vhs_xfwd completes the transform
// along x and accumulates w_p b1_p FFT(S_p) over p into G^.  Same algebra
// and tables as k_vhs_gain_rr (spectrum Ft, phase table, Nyquist planes from
// k_vhs_nyq); only the order of the transforms' stages differs.
// LDS planes are swizzled (column b of row a at nyq_swizzle(b, a)) so the
// row, column and pencil accesses are all conflict free.
// ---------------------------------------------------------------------------
template <int N>
struct SlabShape {
  // slabs per point: as few as the FP64 planes fit in LDS, and a slab size
  // whose FFT is cheap (MI355X, FP64: at 24^3 3 slabs of 8 planes 36.9k Q/s
  // against 22.8k with 2 of 12 and 23.1k with 4 of 6; at 16^3 one slab of 16
  // planes 812k against 537k with 2 of 8)
  static constexpr int NS = (N == 32) ? 4 : (N == 24) ? 3 : 1;
  static constexpr int XS = N / NS;                                              // planes per slab
  static constexpr int NT = XS * N;                                              // one row per thread
  // plane rows padded to N + 1 at 24^3 (row and column accesses conflict
  // free with immediate offsets), else swizzled.  The swizzled offsets, one
  // per element, are held across the item loop by the compiler and cost
  // registers: at 24^3 the padding removed the FP64 kernel's AGPR spills
  // (56.8k -> 59.5k Q/s); at 32^3 it measured slower (97.0 against 95.2 ms
  // per 2048 points in FP64), and at 16^3 it would cost the second
  // workgroup per CU
  static constexpr int RP = (N == 24) ? N + 1 : N;
  static constexpr int PL = N * RP;                                              // complex per plane
  static constexpr int planes = XS * PL;                                         // complex (storage)
  static constexpr int cells = XS * N * N;                                       // values per slab
  static __device__ __forceinline__ int at(int r, int c) {
    if constexpr (RP == N) return r * N + nyq_swizzle<N>(c, r);
    else return r * RP + c;
  }
  static constexpr int wl = N * N + 2 * XS * N;  // complex: W0 plane, the slab's W1 and W2 rows
  static constexpr int wpt = wl / NT;            // Nyquist-plane values per thread
  static_assert(NS * XS == N, "slabs must tile the x axis");
  static_assert(wl % NT == 0, "Nyquist planes split evenly over the workgroup");
  static_assert((N * N) % NT == 0, "x-pass pencils split evenly over the workgroup");
};


template <typename T>
struct GainSlabArgs {
  const cpx<T>* Ft;     // [npts][n] (1 + i) F transposed to [kx][kz][ky]
  const cpx<T>* W;      // [npts][K][3][N][N]
  cpx<T>* Tb;           // [npts][Nrho][N][N][N]: y, z forward transformed slabs of S_p
  const cpx<T>* ephtab; // [K][3][N] phase factors
  int nitems, ngroups, ppg, Nrho, M;
};

// x-pass of slab S (compile-time, so the fold's roots of unity and the
// twiddles are constants: quarter turns cost nothing at NS = 4)
template <typename T, int N, int S>
__device__ __forceinline__ void slab_x_pass(const cpx<T>* __restrict__ Ft, cpx<T>* X, const cpx<T> (*eph)[N]) {
  using SS = SlabShape<N>;
  constexpr int NS = SS::NS, XS = SS::XS, NT = SS::NT, NN = N * N;
  // one pencil per round; the XS spectrum values of fold term m + 1 are in
  // flight while term m is accumulated, and the x phase factors are read
  // from LDS per pencil (an opaque offset keeps the compiler from holding
  // all N of them in registers across the pencils).  Two pencils per round
  // with the factors held in registers left the FP64 kernel so short of
  // registers that every spectrum load was waited for before the next was
  // issued: 143.9 -> 95.0 ms per 2048 points at 32^3 FP64 on MI355X.
#pragma unroll 1
  for (int i = 0; i < NN / NT; ++i) {
    const int idx = static_cast<int>(threadIdx.x) + NT * i;  // kz * N + ky
    const int kz = idx / N, ky = idx - (idx / N) * N;
    const cpx<T> eyz = eph[1][ky] * eph[2][kz];
    const T A = eyz.x + eyz.y, B = eyz.x - eyz.y;  // Re E + Im E = ex.x A + ex.y B
    const cpx<T>* P = Ft + idx;
    int eo = 0;
    asm volatile("" : "+v"(eo));
    const cpx<T>* ex = &eph[0][0] + eo;
    cpx<T> U[XS], ld[2][XS];
#pragma unroll
    for (int r = 0; r < XS; ++r) {
      U[r] = cpx<T>{T(0), T(0)};
      ld[0][r] = P[r * NN];
    }
    static_for<0, NS>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      if constexpr (m + 1 < NS) {
#pragma unroll
        for (int r = 0; r < XS; ++r) ld[(m + 1) & 1][r] = P[(r + XS * (m + 1)) * NN];
      }
      static_for<0, XS>([&](auto rc) {
        constexpr int r = decltype(rc)::value, kx = r + XS * m;
        const cpx<T> e = ex[kx];
        const T fac = e.x * A + e.y * B;
        U[r] = U[r] + FFT<T, NS, +1>::template tw<(m * S) % NS>(fac * ld[m & 1][r]);
      });
    });
    static_for<1, XS>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      U[r] = FFT<T, N, +1>::template tw<(r * S) % N>(U[r]);
    });
    FFT<T, XS, +1, true>::run(U);
#pragma unroll
    for (int j = 0; j < XS; ++j) X[j * SS::PL + SS::at(ky, kz)] = U[j];
  }
}

template <typename T, int N>
__device__ __forceinline__ void slab_x_dispatch(int s, const cpx<T>* __restrict__ Ft, cpx<T>* X,
                                                const cpx<T> (*eph)[N]) {
  constexpr int NS = SlabShape<N>::NS;
  if constexpr (NS == 1) {
    slab_x_pass<T, N, 0>(Ft, X, eph);
  } else if constexpr (NS == 2) {
    if (s == 0) slab_x_pass<T, N, 0>(Ft, X, eph); else slab_x_pass<T, N, 1>(Ft, X, eph);
  } else if constexpr (NS == 3) {
    if (s == 0) slab_x_pass<T, N, 0>(Ft, X, eph);
    else if (s == 1) slab_x_pass<T, N, 1>(Ft, X, eph);
    else slab_x_pass<T, N, 2>(Ft, X, eph);
  } else {
    static_assert(NS == 4, "slab counts 1..4");
    if (s == 0) slab_x_pass<T, N, 0>(Ft, X, eph);
    else if (s == 1) slab_x_pass<T, N, 1>(Ft, X, eph);
    else if (s == 2) slab_x_pass<T, N, 2>(Ft, X, eph);
    else slab_x_pass<T, N, 3>(Ft, X, eph);
  }
}

template <typename T, int N>
__global__ __launch_bounds__(SlabShape<N>::NT) void k_vhs_gain_slab(GainSlabArgs<T> a) {
  using SS = SlabShape<N>;
  constexpr int NS = SS::NS, XS = SS::XS, NN = N * N, n = NN * N;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* W0 = X + SS::planes;  // [y][z]
  cpx<T>* W1 = W0 + NN;         // [j][z]: plane ky = N/2 at x = s + NS j
  cpx<T>* W2 = W1 + XS * N;     // [j][y]: plane kz = N/2 at x = s + NS j
  __shared__ cpx<T> eph[2][3][N];  // phase factors of this (p, q) and the next
  const int tid = threadIdx.x;
  const int j = tid / N, c = tid - (tid / N) * N;  // plane, row (z-pass) / column (y-pass)
  const int K = a.Nrho * a.M;
  T sacc[N];
  for (int item = xcd_slot(); item < a.nitems; item += gridDim.x) {
    const int s = item % NS, pg = item / NS;
    const int pt = pg / a.ngroups, grp = pg - pt * a.ngroups;
    const int p0 = grp * a.ppg, p1 = min(p0 + a.ppg, a.Nrho);
    const int x = s + NS * j;
    const T sx = (x & 1) ? T(-1) : T(1), sz = (c & 1) ? T(-1) : T(1);
    const cpx<T>* Ft = a.Ft + static_cast<size_t>(pt) * n;
    __syncthreads();  // the previous item's readers of eph are done
    if (tid < 3 * N) (&eph[0][0][0])[tid] = a.ephtab[static_cast<size_t>(p0) * a.M * 3 * N + tid];
    int it = 0;  // (p, q) iterations of this item: eph[it & 1] holds the current phases
    for (int p = p0; p < p1; ++p) {
#pragma unroll
      for (int y = 0; y < N; ++y) sacc[y] = T(0);
      for (int q = 0; q < a.M; ++q, ++it) {
        __syncthreads();  // the previous (p, q)'s readers of X and W are done; eph[it & 1] is in place
        // this (p, q)'s Nyquist planes, in flight during the x-pass
        const cpx<T>* Wq = a.W + (static_cast<size_t>(pt) * K + p * a.M + q) * 3 * NN;
        auto wsrc = [&](int e) -> const cpx<T>* {
          if (e < NN) return Wq + e;
          const int e2 = e - NN, pl = e2 / (XS * N), rem = e2 - pl * (XS * N);
          const int jj = rem / N, cc = rem - jj * N;
          return Wq + (1 + pl) * NN + (s + NS * jj) * N + cc;
        };
        if constexpr (sizeof(cpx<T>) == 16) {
          // FP64: the load unit writes them into LDS itself (one gathered
          // 16-byte value per lane, lane i of a wave to the wave's base + 16 i),
          // so they hold no registers across the x-pass
          static_assert(SS::NT % 64 == 0, "whole waves");
          const int wbase = tid & ~63;
#pragma unroll
          for (int k = 0; k < SS::wpt; ++k)
            __builtin_amdgcn_global_load_lds(
                static_cast<const void*>(wsrc(tid + k * SS::NT)),
                (void __attribute__((address_space(3)))*)(W0 + wbase + k * SS::NT), 16, 0, 0);
          slab_x_dispatch<T, N>(s, Ft, X, eph[it & 1]);
          lds_copy_wait();
        } else {
          cpx<T> wv[SS::wpt];
#pragma unroll
          for (int k = 0; k < SS::wpt; ++k) wv[k] = *wsrc(tid + k * SS::NT);
          slab_x_dispatch<T, N>(s, Ft, X, eph[it & 1]);
#pragma unroll
          for (int k = 0; k < SS::wpt; ++k) W0[tid + k * SS::NT] = wv[k];
        }
        __syncthreads();
        {  // z-pass: row (j, ky = c)
          cpx<T>* pl = X + j * SS::PL;
          cpx<T> v[N];
#pragma unroll
          for (int b = 0; b < N; ++b) v[b] = pl[SS::at(c, b)];
          FFT<T, N, +1, true>::run(v);
#pragma unroll
          for (int b = 0; b < N; ++b) pl[SS::at(c, b)] = v[b];
        }
        __syncthreads();
        // the next (p, q)'s phase factors into the other buffer (read after
        // the next iteration's first barrier)
        if (tid < 3 * N) {
          const int qn = (q + 1 < a.M) ? q + 1 : 0, pn = (q + 1 < a.M) ? p : p + 1;
          if (pn < p1) (&eph[(it + 1) & 1][0][0])[tid] = a.ephtab[(static_cast<size_t>(pn) * a.M + qn) * 3 * N + tid];
        }
        {  // y-pass: column (j, z = c), then the products
          const cpx<T>* pl = X + j * SS::PL;
          cpx<T> v[N];
#pragma unroll
          for (int ky = 0; ky < N; ++ky) v[ky] = pl[SS::at(ky, c)];
          FFT<T, N, +1, true>::run(v);
          const cpx<T> w1 = W1[j * N + c];
#pragma unroll
          for (int y = 0; y < N; ++y) {
            const T sy = (y & 1) ? T(-1) : T(1);
            const cpx<T> w0 = W0[y * N + c], w2 = W2[j * N + y];
            const T wr = sx * w0.x + sy * w1.x + sz * w2.x;
            const T wi = sx * w0.y + sy * w1.y + sz * w2.y;
            sacc[y] += (v[y].x + wi) * (v[y].y - wr) - wr * wi;
          }
        }
      }
      // forward transform of the slab of S_p along y (own column) and z
      {
        cpx<T> v[N];
#pragma unroll
        for (int y = 0; y < N; ++y) v[y] = cpx<T>{sacc[y], T(0)};
        FFT<T, N, -1, true>::run(v);
        cpx<T>* pl = X + j * SS::PL;
#pragma unroll
        for (int ky = 0; ky < N; ++ky) pl[SS::at(ky, c)] = v[ky];
      }
      __syncthreads();
      {
        cpx<T>* pl = X + j * SS::PL;
        cpx<T> v[N];
#pragma unroll
        for (int b = 0; b < N; ++b) v[b] = pl[SS::at(c, b)];
        FFT<T, N, -1, true>::run(v);
#pragma unroll
        for (int b = 0; b < N; ++b) pl[SS::at(c, b)] = v[b];
      }
      __syncthreads();
      {
        cpx<T>* Tp = a.Tb + (static_cast<size_t>(pt) * a.Nrho + p) * n;
        for (int e = tid; e < SS::cells; e += SS::NT) {
          const int jj = e / NN, rem = e - jj * NN, ky = rem / N, kz = rem - ky * N;
          Tp[static_cast<size_t>(s + NS * jj) * NN + rem] = X[jj * SS::PL + SS::at(ky, kz)];
        }
      }
    }
  }
}

// x-forward transform of the slab kernel's T and the weighted sum over p:
// G^[pt](k) = sum_p w_p b1_p(k) FFT_x(T_p)(k).  A block of 256 threads takes
// 128 (point, pencil (ky, kz)) columns; waves 2i and 2i + 1 share 64 of them
// and keep the lower and the upper half of the kx sums, so the accumulators
// stay in registers and the half is uniform over each wave.
template <typename T>
struct XfwdArgs {
  const cpx<T>* Tb;     // [npts][Nrho][N][N][N]
  const T* b1;          // [Nrho][n]
  const double* wgt;    // [Nrho]
  cpx<T>* Ghat;         // [npts][n]
  int npts, Nrho;
};

template <typename T, int N, int K0>
__device__ __forceinline__ void xfwd_half(const XfwdArgs<T>& a, int pt, int idx) {
  constexpr int NN = N * N, n = NN * N, H = N / 2;
  cpx<T> acc[H];
#pragma unroll
  for (int k = 0; k < H; ++k) acc[k] = cpx<T>{T(0), T(0)};
  for (int p = 0; p < a.Nrho; ++p) {
    const cpx<T>* Tp = a.Tb + (static_cast<size_t>(pt) * a.Nrho + p) * n + idx;
    cpx<T> v[N];
#pragma unroll
    for (int x = 0; x < N; ++x) v[x] = Tp[static_cast<size_t>(x) * NN];
    FFT<T, N, -1>::run(v);
    const T wp = static_cast<T>(a.wgt[p]);
    const T* b = a.b1 + static_cast<size_t>(p) * n + idx;
#pragma unroll
    for (int k = 0; k < H; ++k) acc[k] = acc[k] + (wp * b[static_cast<size_t>(k + K0) * NN]) * v[k + K0];
  }
  cpx<T>* G = a.Ghat + static_cast<size_t>(pt) * n + idx;
#pragma unroll
  for (int k = 0; k < H; ++k) G[static_cast<size_t>(k + K0) * NN] = acc[k];
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_vhs_xfwd(XfwdArgs<T> a) {
  constexpr int NN = N * N;
  const int w = threadIdx.x / 64, lane = threadIdx.x - (threadIdx.x / 64) * 64;
  const long col = static_cast<long>(blockIdx.x) * 128 + (w >> 1) * 64 + lane;
  if (col >= static_cast<long>(a.npts) * NN) return;
  const int pt = static_cast<int>(col / NN), idx = static_cast<int>(col - static_cast<long>(pt) * NN);
  if (w & 1)
    xfwd_half<T, N, N / 2>(a, pt, idx);
  else
    xfwd_half<T, N, 0>(a, pt, idx);
}

template <typename T>
struct FinArgs {
  const cpx<T>* Ghat;  // [npts*ngroups][n]
  const cpx<T>* F;     // [npts][n]
  const T* b2;         // [n]
  const double* f;     // [npts][n]
  double* Q;           // [npts][n]
  cpx<T>* Z;           // scratch
  double prefac;
  int npts, ngroups;
};

template <typename T, int N>
__global__ __launch_bounds__(Shape<N>::NT) void k_vhs_finalize(FinArgs<T> a) {
  constexpr int NT = Shape<N>::NT, PB = Shape<N>::PB, NN = N * N, n = N * N * N;
  __shared__ cpx<T> lds[PB * N * (N + 1)];
  cpx<T>* Z = a.Z + static_cast<size_t>(blockIdx.x) * n;
  for (int pt = blockIdx.x; pt < a.npts; pt += gridDim.x) {
    const cpx<T>* G = a.Ghat + static_cast<size_t>(pt) * a.ngroups * n;
    const cpx<T>* F = a.F + static_cast<size_t>(pt) * n;
    for (int kx0 = 0; kx0 < N; kx0 += PB) {
      const int cnt = min(PB, N - kx0);
      slab_batch<T, N, NT, +1>(
          lds, cnt,
          [&](int j, int r, int c) -> cpx<T> {
            const int o = (kx0 + j) * NN + r * N + c;
            cpx<T> s = G[o];
            for (int g = 1; g < a.ngroups; ++g) s = s + G[static_cast<size_t>(g) * n + o];
            const cpx<T> lf = a.b2[o] * F[o];
            return s + mul_i(lf);
          },
          [&](int j, int r, int c, cpx<T> v) { Z[(kx0 + j) * NN + r * N + c] = v; });
    }
    const double* f = a.f + static_cast<size_t>(pt) * n;
    double* Q = a.Q + static_cast<size_t>(pt) * n;
    pencil_pass<T, N, NT, +1>([&](int kx, int idx) { return Z[kx * NN + idx]; },
                              [&](int idx, cpx<T>* v) {
                                // the pencil of f is loaded before Q is stored, so the loads are in
                                // flight together (f and Q may overlap as far as the compiler knows)
                                double fv[N];
#pragma unroll
                                for (int x = 0; x < N; ++x) fv[x] = f[x * NN + idx];
#pragma unroll
                                for (int x = 0; x < N; ++x)
                                  Q[x * NN + idx] = a.prefac * (static_cast<double>(v[x].x) -
                                                                fv[x] * static_cast<double>(v[x].y));
                              });
  }
}

// Moments of Q over one point's velocity grid, reduced over the workgroup
// (wave shuffles, then LDS), and the projection's multipliers lambda =
// Ainv (cw m).  m: this thread's partial sums of Q (1, cx, cy, cz, |c|^2);
// red holds NW x 5 doubles; every thread returns the five multipliers.
template <int NW>
__device__ __forceinline__ void block_invariant_multipliers(double (&m)[5], double (*red)[5], double* lam,
                                                            const double* Ainv, double cw) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double sj = m[j];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) sj += __shfl_xor(sj, d);
    m[j] = sj;
  }
  if ((tid & 63) == 0) {
#pragma unroll
    for (int j = 0; j < 5; ++j) red[tid >> 6][j] = m[j];
  }
  __syncthreads();
  if (tid < 5) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      double sk = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) sk += red[w][k];
      acc += Ainv[tid * 5 + k] * (sk * cw);
    }
    lam[tid] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 5; ++j) m[j] = lam[j];
}

// ---------------------------------------------------------------------------
// Register-resident spectrum prep and finalize (FP32, N = 16, 32), in the
// register-resident gain kernel's thread layout: a workgroup of N*N/2
// threads per point, thread t = g N + l holding two pencils.  Neither stages
// the field through a scratch cube in global memory, as the LDS-slab prep
// and finalize do.  Their transforms use the plain butterflies (the
// tangent form's extra rounding is not worth its saving outside the gain).
//   prep:     x-pencils (y = g + (N/2) r, z = l) of f, x-FFT and the kx <-> y
//             exchange (rr_x_forward), then per plane slot kx = g + (N/2) sr
//             y-FFT, in-plane transpose, z-FFT; Ft = (1 + i) F / n, transposed
//             to [kx][kz][ky], is stored from the z-pencils and F / n, after
//             one more in-plane transpose, from the ky-pencils, so both
//             stores are lane-coalesced.
//   finalize: y-pencils (kx, kz = l) of sum_groups G^ + i b2 F (coalesced
//             loads), y-IFFT, transpose, z-IFFT, transpose back to pencils
//             over y at z = l, the gain kernel's exchange to x-pencils
//             (y = g + (N/2) r, z = l), x-IFFT, Q = prefac (Re - f Im).  With
//             PROJ the FP64 projection onto the five collision invariants
//             follows in the same workgroup: Q of both pencils stays in
//             registers while the moments are reduced over the workgroup,
//             so Q is written once, corrected.
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(RRShape<N>::NT) void k_vhs_prep_rr(PrepArgs<T> a) {
  constexpr int G = RRShape<N>::G, NN = N * N, n = NN * N;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* Xb = reinterpret_cast<cpx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  const T rn = static_cast<T>(1.0 / static_cast<double>(n));
  cpx<T> v0[N], v1[N];
  for (int pt = xcd_slot(); pt < a.npts; pt += gridDim.x) {
    const double* f = a.f + static_cast<size_t>(pt) * n;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      T s[N];
      const double* fp = f + (g + G * r) * N + l;
#pragma unroll
      for (int x = 0; x < N; ++x) s[x] = static_cast<T>(fp[x * NN]);
      rr_x_forward<T, N, G, false>(s, v0, v1, Xb, r, g, l);
    }
    cpx<T>* F = a.F + static_cast<size_t>(pt) * n;
    cpx<T>* Ft = a.Ft + static_cast<size_t>(pt) * n;
    __syncthreads();  // the exchange buffer's readers are done before the wave-local transposes reuse it
#pragma unroll
    for (int sr = 0; sr < 2; ++sr) {
      cpx<T>(&v)[N] = sr ? v1 : v0;
      const int kx = g + G * sr;
      FFT<T, N, -1>::run(v);  // y: (kx, ky, z = l)
      __builtin_amdgcn_sched_barrier(0);
      rr_plane_transpose<T, N, G, true, false>(v, Xb, g, l);  // (kx, ky = l, z)
      FFT<T, N, -1>::run(v);                            // z: (kx, ky = l, kz)
      __builtin_amdgcn_sched_barrier(0);
      cpx<T>* ft = Ft + kx * NN + l;
#pragma unroll
      for (int kz = 0; kz < N; ++kz) ft[kz * N] = rn * cpx<T>{v[kz].x - v[kz].y, v[kz].x + v[kz].y};  // (1 + i) F
      rr_plane_transpose<T, N, G, true, true>(v, Xb, g, l);  // (kx, ky, kz = l)
      cpx<T>* fp = F + kx * NN + l;
#pragma unroll
      for (int ky = 0; ky < N; ++ky) fp[ky * N] = rn * v[ky];
    }
  }
}

template <typename T>
struct FinRRArgs {
  FinArgs<T> fa;
  double Ainv[25];  // inverse Gram matrix of the invariants (projection)
  double L, cw;
};

template <typename T, int N, bool PROJ>
__global__ __launch_bounds__(RRShape<N>::NT) void k_vhs_finalize_rr(FinRRArgs<T> ar) {
  constexpr int NT = RRShape<N>::NT, G = RRShape<N>::G, NN = N * N, n = NN * N, NW = NT / 64;
  const FinArgs<T>& a = ar.fa;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* Xb = reinterpret_cast<cpx<T>*>(smem_raw);
  __shared__ double red[NW][5];
  __shared__ double lam[5];
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  const double hvc = 2.0 * ar.L / N;
  cpx<T> v0[N], v1[N];
  for (int pt = xcd_slot(); pt < a.npts; pt += gridDim.x) {
    const cpx<T>* Gp = a.Ghat + static_cast<size_t>(pt) * a.ngroups * n;
    const cpx<T>* F = a.F + static_cast<size_t>(pt) * n;
    // opaque per point, so the compiler does not hoist the b2 addresses and
    // the velocity abscissae of every pencil element out of the point loop
    // (they would hold over a hundred registers)
    int zero = 0;
    asm volatile("" : "+v"(zero));
    const T* b2 = a.b2 + zero;
    const double hv = hvc + static_cast<double>(zero), c0 = -ar.L + 0.5 * hv;
#pragma unroll
    for (int sr = 0; sr < 2; ++sr) {
      cpx<T>(&v)[N] = sr ? v1 : v0;
      const int o0 = (g + G * sr) * NN + l;  // y-pencil (kx, kz = l)
#pragma unroll
      for (int k0 = 0; k0 < N; k0 += 8) {  // 8 values in flight at a time (register pressure)
#pragma unroll
        for (int ky = k0; ky < k0 + 8; ++ky) v[ky] = Gp[o0 + ky * N];
        for (int gr = 1; gr < a.ngroups; ++gr) {  // the other radial groups' partial sums
          const cpx<T>* Gg = Gp + static_cast<size_t>(gr) * n + o0;
#pragma unroll
          for (int ky = k0; ky < k0 + 8; ++ky) v[ky] = v[ky] + Gg[ky * N];
        }
        // the loss term after the whole gain sum (added to one group's
        // partial sum, b2 F is large enough against it to cost precision)
#pragma unroll
        for (int ky = k0; ky < k0 + 8; ++ky) {
          const int o = o0 + ky * N;
          v[ky] = v[ky] + mul_i(b2[o] * F[o]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      FFT<T, N, +1>::run(v);  // y: (kx, y, kz = l)
      __builtin_amdgcn_sched_barrier(0);
      if (sr == 0) __syncthreads();  // the previous point's exchange readers are done
      rr_plane_transpose<T, N, G, true, false>(v, Xb, g, l);  // (kx, y = l, kz)
      FFT<T, N, +1>::run(v);                            // z: (kx, y = l, z)
      __builtin_amdgcn_sched_barrier(0);
      rr_plane_transpose<T, N, G, true, true>(v, Xb, g, l);  // (kx, y, z = l)
    }
    // exchange to the x-pencils of both target slots (v0 and v1 are free
    // after their last writes), then the x-IFFTs
    cpx<T> u0[N], u1[N];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      cpx<T>(&u)[N] = r ? u1 : u0;
      __syncthreads();
      rr_x_write<T, N, G>(v0, Xb, r, g, l);
      __syncthreads();
      rr_x_read<T, N, G>(u, Xb, 0, g, l);
      __syncthreads();
      rr_x_write<T, N, G>(v1, Xb, r, g, l);
      __syncthreads();
      rr_x_read<T, N, G>(u, Xb, 1, g, l);
    }
    FFT<T, N, +1>::run(u0);  // x: (x, y = g, z = l)
    __builtin_amdgcn_sched_barrier(0);
    FFT<T, N, +1>::run(u1);  // x: (x, y = g + G, z = l)
    __builtin_amdgcn_sched_barrier(0);
    const double* f = a.f + static_cast<size_t>(pt) * n;
    double* Q = a.Q + static_cast<size_t>(pt) * n;
    const double cz = c0 + l * hv;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0, l4 = 0.0;
    if constexpr (PROJ) {
      // m = sum over the thread's two pencils of Q phi_j, phi = (1, c, |c|^2);
      // the y and z parts are applied to the pencil sums
      double s0[2] = {0.0, 0.0}, sx = 0.0, sxx = 0.0;
      const double* fp0 = f + g * N + l;
      const double* fp1 = fp0 + G * N;
#pragma unroll
      for (int x0 = 0; x0 < N; x0 += 8) {
#pragma unroll
        for (int x = x0; x < x0 + 8; ++x) {
          const double qa = a.prefac * (static_cast<double>(u0[x].x) - fp0[x * NN] * static_cast<double>(u0[x].y));
          const double qb = a.prefac * (static_cast<double>(u1[x].x) - fp1[x * NN] * static_cast<double>(u1[x].y));
          const double cx = c0 + x * hv, qs = qa + qb;
          s0[0] += qa;
          s0[1] += qb;
          sx += qs * cx;
          sxx += qs * (cx * cx);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      const double cy0 = c0 + g * hv, cy1 = c0 + (g + G) * hv, st = s0[0] + s0[1];
      double m[5] = {st, sx, s0[0] * cy0 + s0[1] * cy1, st * cz,
                     sxx + s0[0] * (cy0 * cy0 + cz * cz) + s0[1] * (cy1 * cy1 + cz * cz)};
      block_invariant_multipliers<NW>(m, red, lam, ar.Ainv, ar.cw);
      l0 = m[0]; l1 = m[1]; l2 = m[2]; l3 = m[3]; l4 = m[4];
    }
    // Q = prefac (Re - f Im), less the projection's correction.  f is
    // re-read for it (L2 hits after the moment pass) through an opaque
    // offset: otherwise the compiler keeps the moment pass's 64 values of Q
    // per thread live across the reduction and spills them
    int zf = 0;
    asm volatile("" : "+v"(zf));
    // likewise the x-pencils and the abscissae: the moment pass's FP64
    // conversions and |c|^2 values are recomputed rather than kept
    double hq = hv;
    asm volatile("" : "+v"(hq));
    const double cq = -ar.L + 0.5 * hq, czq = cq + l * hq;
    if constexpr (PROJ) {
#pragma unroll
      for (int x = 0; x < N; ++x) {
        asm volatile("" : "+v"(u0[x].x), "+v"(u0[x].y));
        asm volatile("" : "+v"(u1[x].x), "+v"(u1[x].y));
      }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const cpx<T>(&u)[N] = r ? u1 : u0;
      const int y = g + G * r;
      const double cy = cq + y * hq;
      const double* fp = f + y * N + l + zf;
      double* qp = Q + y * N + l;
#pragma unroll
      for (int x0 = 0; x0 < N; x0 += 8) {
        double fv[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) fv[x] = fp[(x0 + x) * NN];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
          double q = a.prefac * (static_cast<double>(u[x0 + x].x) - fv[x] * static_cast<double>(u[x0 + x].y));
          if constexpr (PROJ) {
            const double cx = cq + (x0 + x) * hq;
            q -= ar.cw * (l0 + l1 * cx + l2 * cy + l3 * czq + l4 * (cx * cx + cy * cy + czq * czq));
          }
          qp[(x0 + x) * NN] = q;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

// One-pencil register-resident prep and finalize (FP32, N = 24), in the
// one-pencil gain kernel's layout: N*N threads per point, thread t = g N + l,
// the field exchanged through LDS ([kx][y][z], kx stride N (N + 1)) and the
// in-plane transposes through padded kx planes.
//   prep:     x-pencil (y = g, z = l) of f, x-FFT, field exchange to the
//             y-pencil (kx = g, z = l), y-FFT, transpose, z-FFT: Ft stored
//             from the z-pencil, F after one more transpose (both coalesced).
//   finalize: y-pencil (kx = g, kz = l) of sum_groups G^ + i b2 F, y-IFFT,
//             transpose, z-IFFT, transpose to the pencil over y at z = l,
//             field exchange to the x-pencil (y = g, z = l), x-IFFT,
//             Q = prefac (Re - f Im), with PROJ the projection in the kernel.
template <typename T, int N>
__global__ __launch_bounds__(RR1Shape<N>::NT) void k_vhs_prep_rr1(PrepArgs<T> a) {
  constexpr int PP = RR1Shape<N>::PP, KS = RR1Shape<N>::KS, NN = N * N, n = NN * N;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X = reinterpret_cast<cpx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  cpx<T>* Xp = X + g * N * PP;  // kx plane g
  const T rn = static_cast<T>(1.0 / static_cast<double>(n));
  cpx<T> v[N];
  for (int pt = xcd_slot(); pt < a.npts; pt += gridDim.x) {
    const double* fp = a.f + static_cast<size_t>(pt) * n + g * N + l;
#pragma unroll
    for (int x = 0; x < N; ++x) v[x] = cpx<T>{static_cast<T>(fp[x * NN]), T(0)};
    FFT<T, N, -1>::run(v);  // x: (kx, y = g, z = l)
    __syncthreads();        // the previous point's readers of the planes are done
#pragma unroll
    for (int kx = 0; kx < N; ++kx) X[kx * KS + g * N + l] = v[kx];
    __syncthreads();
#pragma unroll
    for (int y = 0; y < N; ++y) v[y] = X[g * KS + y * N + l];
    FFT<T, N, -1>::run(v);  // y: (kx = g, ky, z = l)
    __syncthreads();
#pragma unroll
    for (int ky = 0; ky < N; ++ky) Xp[ky * PP + l] = v[ky];
    __syncthreads();
#pragma unroll
    for (int z = 0; z < N; ++z) v[z] = Xp[l * PP + z];
    FFT<T, N, -1>::run(v);  // z: (kx = g, ky = l, kz)
    cpx<T>* ft = a.Ft + static_cast<size_t>(pt) * n + g * NN + l;
#pragma unroll
    for (int kz = 0; kz < N; ++kz) ft[kz * N] = rn * cpx<T>{v[kz].x - v[kz].y, v[kz].x + v[kz].y};  // (1 + i) F
    __syncthreads();
#pragma unroll
    for (int kz = 0; kz < N; ++kz) Xp[l * PP + kz] = v[kz];
    __syncthreads();
    cpx<T>* fo = a.F + static_cast<size_t>(pt) * n + g * NN + l;
#pragma unroll
    for (int ky = 0; ky < N; ++ky) fo[ky * N] = rn * Xp[ky * PP + l];
  }
}

template <typename T, int N, bool PROJ>
__global__ __launch_bounds__(RR1Shape<N>::NT) void k_vhs_finalize_rr1(FinRRArgs<T> ar) {
  constexpr int NT = RR1Shape<N>::NT, PP = RR1Shape<N>::PP, KS = RR1Shape<N>::KS, NN = N * N, n = NN * N;
  constexpr int NW = (NT + 63) / 64;
  const FinArgs<T>& a = ar.fa;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cpx<T>* X = reinterpret_cast<cpx<T>*>(smem_raw);
  __shared__ double red[NW][5];
  __shared__ double lam[5];
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid - (tid / N) * N;
  cpx<T>* Xp = X + g * N * PP;  // kx plane g
  const double hvc = 2.0 * ar.L / N;
  cpx<T> v[N];
  for (int pt = xcd_slot(); pt < a.npts; pt += gridDim.x) {
    const cpx<T>* Gp = a.Ghat + static_cast<size_t>(pt) * a.ngroups * n + g * NN + l;
    const cpx<T>* F = a.F + static_cast<size_t>(pt) * n + g * NN + l;
    const T* b2 = a.b2 + g * NN + l;
#pragma unroll
    for (int ky = 0; ky < N; ++ky) v[ky] = Gp[ky * N];
    for (int gr = 1; gr < a.ngroups; ++gr) {  // the other radial groups' partial sums
      const cpx<T>* Gg = Gp + static_cast<size_t>(gr) * n;
#pragma unroll
      for (int ky = 0; ky < N; ++ky) v[ky] = v[ky] + Gg[ky * N];
    }
#pragma unroll
    for (int ky = 0; ky < N; ++ky) v[ky] = v[ky] + mul_i(b2[ky * N] * F[ky * N]);  // loss term after the gain sum
    FFT<T, N, +1>::run(v);  // y: (kx = g, y, kz = l)
    __syncthreads();        // the previous point's readers of the planes are done
#pragma unroll
    for (int y = 0; y < N; ++y) Xp[y * PP + l] = v[y];
    __syncthreads();
#pragma unroll
    for (int kz = 0; kz < N; ++kz) v[kz] = Xp[l * PP + kz];
    FFT<T, N, +1>::run(v);  // z: (kx = g, y = l, z)
    __syncthreads();
#pragma unroll
    for (int z = 0; z < N; ++z) Xp[l * PP + z] = v[z];
    __syncthreads();
#pragma unroll
    for (int y = 0; y < N; ++y) v[y] = Xp[y * PP + l];  // (kx = g, y, z = l)
    __syncthreads();
#pragma unroll
    for (int y = 0; y < N; ++y) X[g * KS + y * N + l] = v[y];
    __syncthreads();
#pragma unroll
    for (int kx = 0; kx < N; ++kx) v[kx] = X[kx * KS + g * N + l];
    FFT<T, N, +1>::run(v);  // x: (x, y = g, z = l)
    const double* fp = a.f + static_cast<size_t>(pt) * n + g * N + l;
    double* qp = a.Q + static_cast<size_t>(pt) * n + g * N + l;
    double q[N];
#pragma unroll
    for (int x = 0; x < N; ++x)
      q[x] = a.prefac * (static_cast<double>(v[x].x) - fp[x * NN] * static_cast<double>(v[x].y));
    if constexpr (PROJ) {
      // abscissae from an opaque grid step, so the compiler does not hoist
      // all N of them (and their squares) out of the point loop
      double hv = hvc;
      asm volatile("" : "+v"(hv));
      const double c0 = -ar.L + 0.5 * hv, cy = c0 + g * hv, cz = c0 + l * hv;
      double sq = 0.0, sx = 0.0, sxx = 0.0;
#pragma unroll
      for (int x = 0; x < N; ++x) {
        const double cx = c0 + x * hv;
        sq += q[x];
        sx += q[x] * cx;
        sxx += q[x] * (cx * cx);
      }
      double m[5] = {sq, sx, sq * cy, sq * cz, sxx + sq * (cy * cy + cz * cz)};
      block_invariant_multipliers<NW>(m, red, lam, ar.Ainv, ar.cw);
      asm volatile("" : "+v"(hv));  // recomputed, not kept across the reduction
      const double c0b = -ar.L + 0.5 * hv, cyb = c0b + g * hv, czb = c0b + l * hv;
#pragma unroll
      for (int x = 0; x < N; ++x) {
        const double cx = c0b + x * hv;
        q[x] -= ar.cw * (m[0] + m[1] * cx + m[2] * cyb + m[3] * czb + m[4] * (cx * cx + cyb * cyb + czb * czb));
      }
    }
#pragma unroll
    for (int x = 0; x < N; ++x) qp[x * NN] = q[x];
  }
}

// FP64 projection of Q onto the 5 collision invariants {1, c, |c|^2}: the
// minimum-norm correction d = -C^T (C C^T)^{-1} C Q with C_jr = phi_j(r) cw.
struct ProjArgs {
  double* Q;  // [npts][n]
  double Ainv[25];
  double L, cw;
  int N, npts;
};

__global__ __launch_bounds__(256) void k_vhs_project(ProjArgs a) {
  __shared__ double red[5][256];
  __shared__ double lam[5];
  const int n = a.N * a.N * a.N, tid = threadIdx.x;
  const double h = 2.0 * a.L / a.N, c0 = -a.L + 0.5 * h;
  for (int pt = blockIdx.x; pt < a.npts; pt += gridDim.x) {
    double* Q = a.Q + static_cast<size_t>(pt) * n;
    double m[5] = {0, 0, 0, 0, 0};
    for (int r = tid; r < n; r += 256) {
      const int I = r / (a.N * a.N), J = (r / a.N) % a.N, K = r % a.N;
      const double cx = c0 + I * h, cy = c0 + J * h, cz = c0 + K * h, q = Q[r];
      m[0] += q;
      m[1] += q * cx;
      m[2] += q * cy;
      m[3] += q * cz;
      m[4] += q * (cx * cx + cy * cy + cz * cz);
    }
    for (int j = 0; j < 5; ++j) red[j][tid] = m[j] * a.cw;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s)
        for (int j = 0; j < 5; ++j) red[j][tid] += red[j][tid + s];
      __syncthreads();
    }
    if (tid < 5) {
      double acc = 0;
      for (int k = 0; k < 5; ++k) acc += a.Ainv[tid * 5 + k] * red[k][0];
      lam[tid] = acc;
    }
    __syncthreads();
    for (int r = tid; r < n; r += 256) {
      const int I = r / (a.N * a.N), J = (r / a.N) % a.N, K = r % a.N;
      const double cx = c0 + I * h, cy = c0 + J * h, cz = c0 + K * h;
      Q[r] -= a.cw * (lam[0] + lam[1] * cx + lam[2] * cy + lam[3] * cz +
                      lam[4] * (cx * cx + cy * cy + cz * cz));
    }
    __syncthreads();
  }
}

// AoSoA <-> contiguous [pt][r], pt = (elem - e0) * nupts + upt.  One block
// moves a 32 (r) x 32 (elements of one AoSoA block) tile through LDS so both
// sides are coalesced.
__global__ __launch_bounds__(256) void k_aosoa_gather(const double* __restrict__ in,
                                                      double* __restrict__ out, int nupts, int ldim,
                                                      int vsize, int e0, int ne) {
  __shared__ double tile[32][33];
  const int r0 = blockIdx.x * 32, eb = blockIdx.y, upt = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int elem_t = eb * 32 + tx;
  for (int rr = ty; rr < 32; rr += 8) {
    const int r = r0 + rr;
    if (r < vsize && elem_t >= e0 && elem_t < e0 + ne)
      tile[rr][tx] = in[static_cast<size_t>(upt) * ldim + (static_cast<size_t>(eb) * vsize + r) * 32 + tx];
  }
  __syncthreads();
  for (int ee = ty; ee < 32; ee += 8) {
    const int elem = eb * 32 + ee, r = r0 + tx;
    if (r < vsize && elem >= e0 && elem < e0 + ne)
      out[(static_cast<size_t>(elem - e0) * nupts + upt) * vsize + r] = tile[tx][ee];
  }
}

__global__ __launch_bounds__(256) void k_aosoa_scatter(const double* __restrict__ in,
                                                       double* __restrict__ out, int nupts, int ldim,
                                                       int vsize, int e0, int ne) {
  __shared__ double tile[32][33];
  const int r0 = blockIdx.x * 32, eb = blockIdx.y, upt = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int ee = ty; ee < 32; ee += 8) {
    const int elem = eb * 32 + ee, r = r0 + tx;
    if (r < vsize && elem >= e0 && elem < e0 + ne)
      tile[tx][ee] = in[(static_cast<size_t>(elem - e0) * nupts + upt) * vsize + r];
  }
  __syncthreads();
  const int elem_t = eb * 32 + tx;
  for (int rr = ty; rr < 32; rr += 8) {
    const int r = r0 + rr;
    if (r < vsize && elem_t >= e0 && elem_t < e0 + ne)
      out[static_cast<size_t>(upt) * ldim + (static_cast<size_t>(eb) * vsize + r) * 32 + tx] = tile[rr][tx];
  }
}

int launch_aosoa_gather(const double* f, double* out, int nupts, int ldim, int vsize, int e0, int ne,
                        hipStream_t st) {
  if (ne <= 0 || nupts <= 0) return DGFS_OK;
  const int eb_lo = e0 / 32, eb_hi = (e0 + ne - 1) / 32;
  dim3 grid((vsize + 31) / 32, eb_hi - eb_lo + 1, nupts);
  // the kernel indexes element blocks from the shifted base pointer
  hipLaunchKernelGGL(k_aosoa_gather, grid, dim3(256), 0, st, f + static_cast<size_t>(eb_lo) * vsize * 32,
                     out, nupts, ldim, vsize, e0 - eb_lo * 32, ne);
  DGFS_LAUNCH_CHECK("aosoa_gather");
  return DGFS_OK;
}

int launch_aosoa_scatter(const double* in, double* Q, int nupts, int ldim, int vsize, int e0, int ne,
                         hipStream_t st) {
  if (ne <= 0 || nupts <= 0) return DGFS_OK;
  const int eb_lo = e0 / 32, eb_hi = (e0 + ne - 1) / 32;
  dim3 grid((vsize + 31) / 32, eb_hi - eb_lo + 1, nupts);
  hipLaunchKernelGGL(k_aosoa_scatter, grid, dim3(256), 0, st, in, Q + static_cast<size_t>(eb_lo) * vsize * 32,
                     nupts, ldim, vsize, e0 - eb_lo * 32, ne);
  DGFS_LAUNCH_CHECK("aosoa_scatter");
  return DGFS_OK;
}

// Shared by the other operators of the library: (C C^T)^{-1} of the five
// invariants {1, c, |c|^2} on the Nv^3 grid of [-L, L]^3 with weight cw, and
// the projection launch.
int invariant_gram_inverse(int N, double L, double cw, double* Ainv25) {
  double A[25] = {};
  const double hh = 2.0 * L / N, c0 = -L + 0.5 * hh;
  for (int r = 0; r < N * N * N; ++r) {
    const int I = r / (N * N), J = (r / N) % N, K = r % N;
    const double cx = c0 + I * hh, cy = c0 + J * hh, cz = c0 + K * hh;
    const double phi[5] = {1.0, cx, cy, cz, cx * cx + cy * cy + cz * cz};
    for (int j = 0; j < 5; ++j)
      for (int k = 0; k < 5; ++k) A[j * 5 + k] += phi[j] * phi[k] * cw * cw;
  }
  double m[5][10];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 10; ++j) m[i][j] = (j < 5) ? A[i * 5 + j] : (j - 5 == i ? 1.0 : 0.0);
  for (int c = 0; c < 5; ++c) {
    int piv = c;
    for (int r = c + 1; r < 5; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
    if (std::fabs(m[piv][c]) < 1e-300) return fail(DGFS_ERR_INVALID_VALUE, "singular invariant Gram matrix");
    for (int j = 0; j < 10; ++j) std::swap(m[c][j], m[piv][j]);
    const double d = m[c][c];
    for (int j = 0; j < 10; ++j) m[c][j] /= d;
    for (int r = 0; r < 5; ++r)
      if (r != c) {
        const double fct = m[r][c];
        for (int j = 0; j < 10; ++j) m[r][j] -= fct * m[c][j];
      }
  }
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) Ainv25[i * 5 + j] = m[i][j + 5];
  return DGFS_OK;
}

int launch_invariant_projection(double* Q, const double* Ainv25, double L, double cw, int N, int npts,
                                hipStream_t st) {
  if (npts <= 0) return DGFS_OK;
  ProjArgs pr;
  pr.Q = Q;
  std::memcpy(pr.Ainv, Ainv25, sizeof(pr.Ainv));
  pr.L = L; pr.cw = cw; pr.N = N; pr.npts = npts;
  hipLaunchKernelGGL(k_vhs_project, dim3(std::min(npts, 4 * cu_count())), dim3(256), 0, st, pr);
  DGFS_LAUNCH_CHECK("invariant_projection");
  return DGFS_OK;
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
struct VhsHandle {
  int magic = 0x76687331;
  int N = 0, n = 0, Nrho = 0, M = 0, prec = 64, project = 0;
  double L = 0, gamma = 0, prefac = 0, sw = 0, cw = 0;
  size_t tsize = 8;  // sizeof(T)
  void* b1 = nullptr;
  void* b2 = nullptr;
  double* alpha = nullptr;
  double* wgt = nullptr;
  double* szd = nullptr;
  void* ephtab = nullptr;  // [K][3][N] cpx<T>: exp(i alpha_p sz_qd l(k)), per (p, q), axis d, index k
  double Ainv[25] = {};
  // workspace
  size_t cap_pts = 0, cap_items = 0;
  int cap_slots = 0;
  double* fc = nullptr;  // [cap_pts][n]
  double* Qc = nullptr;  // [cap_pts][n]
  void* F = nullptr;     // [cap_pts][n] cpx<T>
  void* Ghat = nullptr;  // [cap_items][n] cpx<T>
  void* Z = nullptr;
  void* S = nullptr;
  void* W = nullptr;
  int groups_override = 0;
  int slots_override = 0;  // cap on the persistent multi-pass grid (0: occupancy)
  int algorithm = 0;     // 0 automatic, 1 multi-pass, 2 register-resident (Nyquist planes
                         // from k_vhs_nyq), 3 register-resident with in-kernel Nyquist planes,
                         // 4 register-resident with one pencil per thread (Nyquist planes
                         // from k_vhs_nyq), 5 as 2 with wave-local in-plane transposes,
                         // 6 slab kernel (x-slabs in LDS, any precision, Nv 16/24/32)
  int last_algorithm = 0;
  void* Wny = nullptr;   // Nyquist planes [cap_pts][K][3][N][N]
  void* Ft = nullptr;    // spectrum transposed to [kx][kz][ky] [cap_pts][n]
  size_t cap_ft = 0;
  size_t cap_wny = 0;
  void* Tb = nullptr;    // slab kernel's forward-transformed slabs [cap_pts][Nrho][n]
  size_t cap_tb = 0;
  int last_groups = 0, last_slots = 0, threads = 0;
  // set for the duration of an allocation-free call: the workspace pointers
  // above then point into the device scratch array and growing any of them
  // is an error instead of a hipMalloc
  int no_alloc = 0;
  // this device's scratch array address (resolved at create) or null, with
  // the failure text when it could not be resolved
  void* scratch = nullptr;
  std::string scratch_err;
  // optional per-kernel timing: event pairs recorded around every launch on
  // the launch stream, resolved (and synchronised) only when queried
  int timing = 0;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double acc_ms[5] = {0, 0, 0, 0, 0};
  int acc_n[5] = {0, 0, 0, 0, 0};
};


static void free_ws(VhsHandle* h) {
  (void)hipFree(h->Wny);
  h->Wny = nullptr;
  h->cap_wny = 0;
  (void)hipFree(h->Tb);
  h->Tb = nullptr;
  h->cap_tb = 0;
  (void)hipFree(h->Ft);
  h->Ft = nullptr;
  h->cap_ft = 0;
  (void)hipFree(h->fc); (void)hipFree(h->Qc); (void)hipFree(h->F); (void)hipFree(h->Ghat);
  (void)hipFree(h->Z); (void)hipFree(h->S); (void)hipFree(h->W);
  h->fc = h->Qc = nullptr; h->F = h->Ghat = h->Z = h->S = h->W = nullptr;
  h->cap_pts = h->cap_items = 0; h->cap_slots = 0;
}

static int invert5(const double* A, double* Ainv) {
  double m[5][10];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 10; ++j) m[i][j] = (j < 5) ? A[i * 5 + j] : (j - 5 == i ? 1.0 : 0.0);
  for (int c = 0; c < 5; ++c) {
    int piv = c;
    for (int r = c + 1; r < 5; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
    if (std::fabs(m[piv][c]) < 1e-300) return 1;
    for (int j = 0; j < 10; ++j) std::swap(m[c][j], m[piv][j]);
    const double d = m[c][c];
    for (int j = 0; j < 10; ++j) m[c][j] /= d;
    for (int r = 0; r < 5; ++r)
      if (r != c) {
        const double fct = m[r][c];
        for (int j = 0; j < 10; ++j) m[r][j] -= fct * m[c][j];
      }
  }
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) Ainv[i * 5 + j] = m[i][j + 5];
  return 0;
}

static bool supported_N(int N) {
  switch (N) {
    case 4: case 6: case 8: case 12: case 16: case 24: case 32: return true;
    default: return false;
  }
}

template <typename T, int N>
static int occupancy_slots() {
  int occ = 1;
  int o1 = 1, o2 = 1, o3 = 1;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o1, reinterpret_cast<const void*>(k_vhs_gain<T, N>), Shape<N>::NT, 0);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o2, reinterpret_cast<const void*>(k_vhs_prep<T, N>), Shape<N>::NT, 0);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o3, reinterpret_cast<const void*>(k_vhs_finalize<T, N>), Shape<N>::NT, 0);
  occ = std::max(1, std::min(o1, std::min(o2, o3)));
  return occ * cu_count();
}

// resident workgroups of a register-resident prep or finalize kernel over
// the device (the persistent grid size)
static int rr_edge_grid(const void* kfn, int threads, size_t lds, int npts) {
  int o = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, kfn, threads, lds) != hipSuccess || o < 1) o = 1;
  return std::max(1, std::min(npts, o * cu_count()));
}

// the register-resident gain kernels exist for FP32: two pencils per thread
// at N in {16, 32}, one pencil per thread at N in {16, 24}
template <typename T, int N>
constexpr bool rr_available() {
  return std::is_same<T, float>::value && (N == 16 || N == 32);
}
template <typename T, int N>
constexpr bool rr1_available() {
  return std::is_same<T, float>::value && (N == 16 || N == 24);
}
static bool rr2_ok(const VhsHandle* h) { return h->prec == 32 && (h->N == 16 || h->N == 32); }
static bool rr1_ok(const VhsHandle* h) { return h->prec == 32 && (h->N == 16 || h->N == 24); }
static bool slab_ok(const VhsHandle* h) { return h->N == 16 || h->N == 24 || h->N == 32; }
template <int N>
constexpr bool slab_available() {
  return N == 16 || N == 24 || N == 32;
}

// gain algorithm a launch uses: 1 multi-pass, 2/3/5 two pencils per thread,
// 4 one pencil per thread, 6 slab kernel.  The automatic choice takes, in
// FP32, the one-pencil kernel where the field fits in LDS (Nv 16, 24: 1.8x
// the two-pencil kernel at 16^3 on MI355X) and the two-pencil kernel with
// wave-local in-plane transposes at Nv 32 (61.0 against 62.0 ms per
// 4096-point launch with the workgroup-wide ones, bitwise the same Q); in
// FP64 the slab kernel at Nv 16, 24, 32 (2.3x, 1.8x, 2.7x the multi-pass
// kernel at 16^3, 24^3, 32^3); the multi-pass one otherwise.
static int rr_kind(const VhsHandle* h) {
  switch (h->algorithm) {
    case 1: return 1;
    case 2: return rr2_ok(h) ? 2 : 1;
    case 3: return rr2_ok(h) ? 3 : 1;
    case 4: return rr1_ok(h) ? 4 : 1;
    case 5: return rr2_ok(h) ? 5 : 1;
    case 6: return slab_ok(h) ? 6 : 1;
    default: return rr1_ok(h) ? 4 : (rr2_ok(h) ? 5 : (slab_ok(h) ? 6 : 1));
  }
}

static bool use_rr(const VhsHandle* h) { return rr_kind(h) != 1; }

template <typename T, int N>
static int run_gain_rr(VhsHandle* h, int npts, int groups, int ppg, size_t items, hipStream_t st) {
  if constexpr (rr_available<T, N>()) {
    constexpr int NN = N * N;
    const int K = h->Nrho * h->M;
    const int kind = rr_kind(h);
    const bool inl = (kind == 3);
    h->last_algorithm = kind;
    const size_t wbytes = static_cast<size_t>(npts) * K * 3 * NN * sizeof(cpx<T>);
    if (!inl && wbytes > h->cap_wny) {
      if (h->no_alloc) return fail(DGFS_ERR_ALLOC, "allocation-free call: Nyquist planes exceed their scratch share");
      (void)hipFree(h->Wny);
      h->Wny = nullptr;
      DGFS_HIP_TRY(hipMalloc(&h->Wny, wbytes));
      h->cap_wny = wbytes;
    }
    if (!inl) {
    const long nplanes = static_cast<long>(npts) * K * 3;
    NyqArgs<T> na{static_cast<const cpx<T>*>(h->F), static_cast<const cpx<T>*>(h->Ft),
                  static_cast<cpx<T>*>(h->Wny), static_cast<const cpx<T>*>(h->ephtab), nplanes, h->Nrho, h->M};
    const long waves = (nplanes + NyqShape<N>::PW - 1) / NyqShape<N>::PW;
    KTimerT<VhsHandle> t4(h, 4, st);
    hipLaunchKernelGGL((k_vhs_nyq<T, N>), dim3(static_cast<unsigned>(std::min<long>(waves, 16L * cu_count()))),
                       dim3(64), 0, st, na);
    DGFS_LAUNCH_CHECK("vhs_nyq");
    t4.stop();
    }
    KTimerT<VhsHandle> t1(h, 1, st);
    const size_t lds = (static_cast<size_t>(RRLds<N>::xsize) + RRLds<N>::wl) * sizeof(cpx<T>) +
                       static_cast<size_t>(RRLds<N>::s0) * sizeof(T);
    const void* kfn = inl ? reinterpret_cast<const void*>(k_vhs_gain_rr<T, N, true>)
                      : (kind == 5) ? reinterpret_cast<const void*>(k_vhs_gain_rr<T, N, false, true>)
                                    : reinterpret_cast<const void*>(k_vhs_gain_rr<T, N, false>);
    DGFS_HIP_TRY(hipFuncSetAttribute(kfn,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const int grid = static_cast<int>(std::min<size_t>(items, static_cast<size_t>(cu_count())));
    GainRRArgs<T> ga{static_cast<const cpx<T>*>(h->Ft), static_cast<const cpx<T>*>(h->Wny),
                     static_cast<cpx<T>*>(h->Ghat), static_cast<const T*>(h->b1), h->alpha, h->wgt, h->szd,
                     static_cast<const cpx<T>*>(h->ephtab), static_cast<int>(items), groups, ppg, h->Nrho, h->M};
    if (inl)
      hipLaunchKernelGGL((k_vhs_gain_rr<T, N, true>), dim3(grid), dim3(RRShape<N>::NT), lds, st, ga);
    else if (kind == 5)
      hipLaunchKernelGGL((k_vhs_gain_rr<T, N, false, true>), dim3(grid), dim3(RRShape<N>::NT), lds, st, ga);
    else
      hipLaunchKernelGGL((k_vhs_gain_rr<T, N, false>), dim3(grid), dim3(RRShape<N>::NT), lds, st, ga);
    DGFS_LAUNCH_CHECK("vhs_gain_rr");
    t1.stop();
    h->threads = RRShape<N>::NT;
    h->last_slots = grid;
    return DGFS_OK;
  } else {
    (void)h; (void)npts; (void)groups; (void)ppg; (void)items; (void)st;
    return fail(DGFS_ERR_UNSUPPORTED, "register-resident kernel not built for this precision/N");
  }
}

template <typename T, int N>
static int run_gain_rr1(VhsHandle* h, int npts, int groups, int ppg, size_t items, hipStream_t st) {
  if constexpr (rr1_available<T, N>()) {
    constexpr int NN = N * N;
    const int K = h->Nrho * h->M;
    h->last_algorithm = 4;
    const size_t wbytes = static_cast<size_t>(npts) * K * 3 * NN * sizeof(cpx<T>);
    if (wbytes > h->cap_wny) {
      if (h->no_alloc) return fail(DGFS_ERR_ALLOC, "allocation-free call: Nyquist planes exceed their scratch share");
      (void)hipFree(h->Wny);
      h->Wny = nullptr;
      DGFS_HIP_TRY(hipMalloc(&h->Wny, wbytes));
      h->cap_wny = wbytes;
    }
    const long nplanes = static_cast<long>(npts) * K * 3;
    NyqArgs<T> na{static_cast<const cpx<T>*>(h->F), static_cast<const cpx<T>*>(h->Ft),
                  static_cast<cpx<T>*>(h->Wny), static_cast<const cpx<T>*>(h->ephtab), nplanes, h->Nrho, h->M};
    const long waves = (nplanes + NyqShape<N>::PW - 1) / NyqShape<N>::PW;
    KTimerT<VhsHandle> t4(h, 4, st);
    hipLaunchKernelGGL((k_vhs_nyq<T, N>), dim3(static_cast<unsigned>(std::min<long>(waves, 16L * cu_count()))),
                       dim3(64), 0, st, na);
    DGFS_LAUNCH_CHECK("vhs_nyq");
    t4.stop();
    KTimerT<VhsHandle> t1(h, 1, st);
    const size_t lds = (static_cast<size_t>(RR1Shape<N>::field) + RR1Shape<N>::wl) * sizeof(cpx<T>);
    const void* kfn = reinterpret_cast<const void*>(k_vhs_gain_rr1<T, N>);
    DGFS_HIP_TRY(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, RR1Shape<N>::NT, lds) != hipSuccess || occ < 1)
      occ = 1;
    const int grid = static_cast<int>(std::min<size_t>(items, static_cast<size_t>(occ) * cu_count()));
    GainRRArgs<T> ga{static_cast<const cpx<T>*>(h->Ft), static_cast<const cpx<T>*>(h->Wny),
                     static_cast<cpx<T>*>(h->Ghat), static_cast<const T*>(h->b1), h->alpha, h->wgt, h->szd,
                     static_cast<const cpx<T>*>(h->ephtab), static_cast<int>(items), groups, ppg, h->Nrho, h->M};
    hipLaunchKernelGGL((k_vhs_gain_rr1<T, N>), dim3(grid), dim3(RR1Shape<N>::NT), lds, st, ga);
    DGFS_LAUNCH_CHECK("vhs_gain_rr1");
    t1.stop();
    h->threads = RR1Shape<N>::NT;
    h->last_slots = grid;
    return DGFS_OK;
  } else {
    (void)h; (void)npts; (void)groups; (void)ppg; (void)items; (void)st;
    return fail(DGFS_ERR_UNSUPPORTED, "one-pencil register-resident kernel not built for this precision/N");
  }
}

// The slab kernel's items of one point read the same spectrum once per
// (p, q); at least 4 radial groups keep 16 of them (4 slabs x 4 groups) per
// point, so the workgroups of an XCD share the spectra of two points in its L2
// (32^3 FP64 on MI355X: 229.7 ms per 2048 points with 1 group, 157.4 with 2,
// 156.1 with 4; 16^3: 2.3 against 1.8 ms per 1024 points)
constexpr int kSlabMinGroups = 4;

// slab gain: Nyquist planes, the slab kernel over (point, group, slab) items,
// then the x-forward sum into G^ (one partial per point)
template <typename T, int N>
static int run_gain_slab(VhsHandle* h, int npts, int groups, int ppg, hipStream_t st) {
  if constexpr (slab_available<N>()) {
    using SS = SlabShape<N>;
    constexpr int NN = N * N;
    constexpr size_t n = static_cast<size_t>(NN) * N;
    const int K = h->Nrho * h->M;
    h->last_algorithm = 6;
    const size_t wbytes = static_cast<size_t>(npts) * K * 3 * NN * sizeof(cpx<T>);
    if (wbytes > h->cap_wny) {
      if (h->no_alloc) return fail(DGFS_ERR_ALLOC, "allocation-free call: Nyquist planes exceed their scratch share");
      (void)hipFree(h->Wny);
      h->Wny = nullptr;
      h->cap_wny = 0;
      DGFS_HIP_TRY(hipMalloc(&h->Wny, wbytes));
      h->cap_wny = wbytes;
    }
    const size_t tbytes = static_cast<size_t>(npts) * h->Nrho * n * sizeof(cpx<T>);
    if (tbytes > h->cap_tb) {
      if (h->no_alloc) return fail(DGFS_ERR_ALLOC, "allocation-free call: slab transforms exceed their scratch share");
      (void)hipFree(h->Tb);
      h->Tb = nullptr;
      h->cap_tb = 0;
      DGFS_HIP_TRY(hipMalloc(&h->Tb, tbytes));
      h->cap_tb = tbytes;
    }
    const long nplanes = static_cast<long>(npts) * K * 3;
    NyqArgs<T> na{static_cast<const cpx<T>*>(h->F), static_cast<const cpx<T>*>(h->Ft),
                  static_cast<cpx<T>*>(h->Wny), static_cast<const cpx<T>*>(h->ephtab), nplanes, h->Nrho, h->M};
    const long waves = (nplanes + NyqShape<N>::PW - 1) / NyqShape<N>::PW;
    {
      KTimerT<VhsHandle> t4(h, 4, st);
      hipLaunchKernelGGL((k_vhs_nyq<T, N>), dim3(static_cast<unsigned>(std::min<long>(waves, 16L * cu_count()))),
                         dim3(64), 0, st, na);
      DGFS_LAUNCH_CHECK("vhs_nyq");
      t4.stop();
    }
    const size_t items = static_cast<size_t>(npts) * groups * SS::NS;
    const size_t lds = (static_cast<size_t>(SS::planes) + SS::wl) * sizeof(cpx<T>);
    const void* kfn = reinterpret_cast<const void*>(k_vhs_gain_slab<T, N>);
    DGFS_HIP_TRY(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, SS::NT, lds) != hipSuccess || occ < 1) occ = 1;
    const int grid = static_cast<int>(std::min<size_t>(items, static_cast<size_t>(occ) * cu_count()));
    GainSlabArgs<T> ga{static_cast<const cpx<T>*>(h->Ft), static_cast<const cpx<T>*>(h->Wny),
                       static_cast<cpx<T>*>(h->Tb), static_cast<const cpx<T>*>(h->ephtab),
                       static_cast<int>(items), groups, ppg, h->Nrho, h->M};
    KTimerT<VhsHandle> t1(h, 1, st);
    hipLaunchKernelGGL((k_vhs_gain_slab<T, N>), dim3(grid), dim3(SS::NT), lds, st, ga);
    DGFS_LAUNCH_CHECK("vhs_gain_slab");
    XfwdArgs<T> xa{static_cast<const cpx<T>*>(h->Tb), static_cast<const T*>(h->b1), h->wgt,
                   static_cast<cpx<T>*>(h->Ghat), npts, h->Nrho};
    const long cols = static_cast<long>(npts) * NN;
    hipLaunchKernelGGL((k_vhs_xfwd<T, N>), dim3(static_cast<unsigned>((cols + 127) / 128)), dim3(256), 0, st, xa);
    DGFS_LAUNCH_CHECK("vhs_xfwd");
    t1.stop();
    h->threads = SS::NT;
    h->last_slots = grid;
    return DGFS_OK;
  } else {
    (void)h; (void)npts; (void)groups; (void)ppg; (void)st;
    return fail(DGFS_ERR_UNSUPPORTED, "slab gain kernel not built for Nv=%d", N);
  }
}

// radial groups per point: enough (point, group) items for two per slot
// radial groups per point.  Automatic: the split whose (point, group) items
// fill the persistent grid in the fewest rounds of radial work: rounds of
// items ceil(npts g / slots), each ppg = ceil(Nrho / g) radial nodes long, with
// every group beyond the first costing 1 % (the finalize kernel reads its
// partial G^).  A batch that is not a multiple of the grid then splits finer
// rather than leaving most of its last round idle (3600 points on 256 CUs:
// 2 groups, 14.5 rounds of full items instead of 15).
static int pick_groups(const VhsHandle* h, int npts, int slots, int min_groups = 1) {
  int groups = h->groups_override;
  if (groups <= 0) {
    const long P = std::max(npts, 1), S = std::max(slots, 1);
    const int g0 = std::max(1, std::min(min_groups, h->Nrho));
    double best = 0.0;
    groups = g0;
    for (int g = g0; g <= h->Nrho; ++g) {
      const int ppg = (h->Nrho + g - 1) / g, ge = (h->Nrho + ppg - 1) / ppg;
      if (ge != g) continue;  // the same split as a smaller g
      const double cost = static_cast<double>((P * ge + S - 1) / S) * ppg * (1.0 + 0.01 * (ge - 1));
      if (g == g0 || cost < best) {
        best = cost;
        groups = ge;
      }
    }
  }
  const int ppg = (h->Nrho + groups - 1) / groups;
  return (h->Nrho + ppg - 1) / ppg;
}

// spectra of the chunk (and the transposed copy for the register-resident
// gain), gain partial buffer; returns the gain kernel's arguments
template <typename T, int N>
static int stage_prep(VhsHandle* h, const double* fc, int npts, int groups, int slots, bool rr, GainArgs<T>* ga,
                      hipStream_t st) {
  constexpr int n = N * N * N;
  const int ppg = (h->Nrho + groups - 1) / groups;
  const size_t items = static_cast<size_t>(npts) * groups;
  if (items > h->cap_items) {
    if (h->no_alloc) return fail(DGFS_ERR_ALLOC, "allocation-free call: gain partials exceed their scratch share");
    (void)hipFree(h->Ghat);
    h->Ghat = nullptr;
    h->cap_items = 0;
    DGFS_HIP_TRY(hipMalloc(&h->Ghat, items * n * sizeof(cpx<T>)));
    h->cap_items = items;
  }
  h->last_groups = groups;
  h->last_slots = slots;
  h->threads = Shape<N>::NT;
  if (rr) {
    const size_t fbytes = static_cast<size_t>(npts) * n * sizeof(cpx<T>);
    if (fbytes > h->cap_ft) {
      if (h->no_alloc) return fail(DGFS_ERR_ALLOC, "allocation-free call: transposed spectra exceed their scratch share");
      (void)hipFree(h->Ft);
      h->Ft = nullptr;
      h->cap_ft = 0;
      DGFS_HIP_TRY(hipMalloc(&h->Ft, fbytes));
      h->cap_ft = fbytes;
    }
  }
  PrepArgs<T> pa{fc, static_cast<cpx<T>*>(h->F), rr ? static_cast<cpx<T>*>(h->Ft) : nullptr,
                 static_cast<cpx<T>*>(h->Z), npts};
  KTimerT<VhsHandle> t0(h, 0, st);
  bool done = false;
  if constexpr (rr_available<T, N>()) {
    if (rr) {  // register-resident prep: F and Ft without the scratch cube
      const size_t lds = static_cast<size_t>(RRLds<N>::xsize) * sizeof(cpx<T>);
      DGFS_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_vhs_prep_rr<T, N>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      const int grid = rr_edge_grid(reinterpret_cast<const void*>(k_vhs_prep_rr<T, N>), RRShape<N>::NT, lds, npts);
      hipLaunchKernelGGL((k_vhs_prep_rr<T, N>), dim3(grid), dim3(RRShape<N>::NT), lds, st, pa);
      DGFS_LAUNCH_CHECK("vhs_prep_rr");
      done = true;
    }
  } else if constexpr (rr1_available<T, N>()) {
    if (rr) {  // one-pencil register-resident prep
      const size_t lds = static_cast<size_t>(RR1Shape<N>::field) * sizeof(cpx<T>);
      DGFS_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_vhs_prep_rr1<T, N>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      const int grid = rr_edge_grid(reinterpret_cast<const void*>(k_vhs_prep_rr1<T, N>), RR1Shape<N>::NT, lds, npts);
      hipLaunchKernelGGL((k_vhs_prep_rr1<T, N>), dim3(grid), dim3(RR1Shape<N>::NT), lds, st, pa);
      DGFS_LAUNCH_CHECK("vhs_prep_rr1");
      done = true;
    }
  }
  if (!done) {
    hipLaunchKernelGGL((k_vhs_prep<T, N>), dim3(std::min(npts, h->cap_slots)), dim3(Shape<N>::NT), 0, st, pa);
    DGFS_LAUNCH_CHECK("vhs_prep");
  }
  t0.stop();
  *ga = GainArgs<T>{static_cast<const cpx<T>*>(h->F), static_cast<cpx<T>*>(h->Ghat),
                    static_cast<cpx<T>*>(h->Z), static_cast<T*>(h->S), static_cast<cpx<T>*>(h->W),
                    static_cast<const T*>(h->b1), h->alpha, h->wgt, h->szd,
                    static_cast<int>(items), groups, ppg, h->Nrho, h->M};
  return DGFS_OK;
}

// gain + loss assembly and the optional FP64 projection
template <typename T, int N>
static int stage_finish(VhsHandle* h, const double* fc, double* Qc, int npts, int groups, hipStream_t st) {
  FinArgs<T> fa{static_cast<const cpx<T>*>(h->Ghat), static_cast<const cpx<T>*>(h->F),
                static_cast<const T*>(h->b2), fc, Qc, static_cast<cpx<T>*>(h->Z),
                h->prefac, npts, groups};
  if constexpr (rr_available<T, N>()) {
    // register-resident finalize, with the projection in the same kernel
    FinRRArgs<T> fr;
    fr.fa = fa;
    std::memcpy(fr.Ainv, h->Ainv, sizeof(fr.Ainv));
    fr.L = h->L;
    fr.cw = h->cw;
    const size_t lds = static_cast<size_t>(RRLds<N>::xsize) * sizeof(cpx<T>);
    const void* kfn = h->project ? reinterpret_cast<const void*>(k_vhs_finalize_rr<T, N, true>)
                                 : reinterpret_cast<const void*>(k_vhs_finalize_rr<T, N, false>);
    DGFS_HIP_TRY(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const dim3 grid(rr_edge_grid(kfn, RRShape<N>::NT, lds, npts));
    KTimerT<VhsHandle> t2(h, 2, st);
    if (h->project)
      hipLaunchKernelGGL((k_vhs_finalize_rr<T, N, true>), grid, dim3(RRShape<N>::NT), lds, st, fr);
    else
      hipLaunchKernelGGL((k_vhs_finalize_rr<T, N, false>), grid, dim3(RRShape<N>::NT), lds, st, fr);
    DGFS_LAUNCH_CHECK("vhs_finalize_rr");
    t2.stop();
    return DGFS_OK;
  } else if constexpr (rr1_available<T, N>()) {
    FinRRArgs<T> fr;
    fr.fa = fa;
    std::memcpy(fr.Ainv, h->Ainv, sizeof(fr.Ainv));
    fr.L = h->L;
    fr.cw = h->cw;
    const size_t lds = static_cast<size_t>(RR1Shape<N>::field) * sizeof(cpx<T>);
    const void* kfn = h->project ? reinterpret_cast<const void*>(k_vhs_finalize_rr1<T, N, true>)
                                 : reinterpret_cast<const void*>(k_vhs_finalize_rr1<T, N, false>);
    DGFS_HIP_TRY(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const dim3 grid(rr_edge_grid(kfn, RR1Shape<N>::NT, lds, npts));
    KTimerT<VhsHandle> t2(h, 2, st);
    if (h->project)
      hipLaunchKernelGGL((k_vhs_finalize_rr1<T, N, true>), grid, dim3(RR1Shape<N>::NT), lds, st, fr);
    else
      hipLaunchKernelGGL((k_vhs_finalize_rr1<T, N, false>), grid, dim3(RR1Shape<N>::NT), lds, st, fr);
    DGFS_LAUNCH_CHECK("vhs_finalize_rr1");
    t2.stop();
    return DGFS_OK;
  }
  KTimerT<VhsHandle> t2(h, 2, st);
  hipLaunchKernelGGL((k_vhs_finalize<T, N>), dim3(std::min(npts, h->cap_slots)), dim3(Shape<N>::NT), 0, st, fa);
  DGFS_LAUNCH_CHECK("vhs_finalize");
  t2.stop();
  if (h->project) {
    ProjArgs pr;
    pr.Q = Qc;
    std::memcpy(pr.Ainv, h->Ainv, sizeof(pr.Ainv));
    pr.L = h->L; pr.cw = h->cw; pr.N = N; pr.npts = npts;
    KTimerT<VhsHandle> t3(h, 3, st);
    hipLaunchKernelGGL(k_vhs_project, dim3(std::min(npts, 4 * cu_count())), dim3(256), 0, st, pr);
    DGFS_LAUNCH_CHECK("vhs_project");
    t3.stop();
  }
  return DGFS_OK;
}

template <typename T, int N>
static int run_chunk(VhsHandle* h, const double* fc, double* Qc, int npts, hipStream_t st) {
  const bool rr = use_rr(h);
  int slots = rr ? cu_count() : h->cap_slots;
  if (!rr && h->slots_override > 0) slots = std::min(slots, h->slots_override);
  if constexpr (slab_available<N>()) {
    if (rr_kind(h) == 6) {  // items are (point, group, slab): each slab counts as a point
      const int groups = pick_groups(h, npts * SlabShape<N>::NS, slots, kSlabMinGroups);
      GainArgs<T> ga;
      int rc = stage_prep<T, N>(h, fc, npts, groups, slots, true, &ga, st);
      if (!rc) rc = run_gain_slab<T, N>(h, npts, groups, ga.ppg, st);
      if (!rc) rc = stage_finish<T, N>(h, fc, Qc, npts, 1, st);
      return rc;
    }
  }
  const int groups = pick_groups(h, npts, slots);
  GainArgs<T> ga;
  int rc = stage_prep<T, N>(h, fc, npts, groups, slots, rr, &ga, st);
  if (rc) return rc;
  if (rr) {
    if (rr_kind(h) == 4)
      rc = run_gain_rr1<T, N>(h, npts, groups, ga.ppg, static_cast<size_t>(ga.nitems), st);
    else
      rc = run_gain_rr<T, N>(h, npts, groups, ga.ppg, static_cast<size_t>(ga.nitems), st);
    if (rc) return rc;
  } else {
    h->last_algorithm = 1;
    KTimerT<VhsHandle> t1(h, 1, st);
    hipLaunchKernelGGL((k_vhs_gain<T, N>), dim3(std::min(ga.nitems, slots)), dim3(Shape<N>::NT), 0, st, ga);
    DGFS_LAUNCH_CHECK("vhs_gain");
    t1.stop();
  }
  return stage_finish<T, N>(h, fc, Qc, npts, groups, st);
}

template <typename T, int N>
static int ensure_ws(VhsHandle* h, int pts) {
  constexpr int n = N * N * N;
  if (h->no_alloc && (h->cap_slots == 0 || static_cast<size_t>(pts) > h->cap_pts))
    return fail(DGFS_ERR_ALLOC, "allocation-free call: point buffers exceed their scratch share");
  if (h->cap_slots == 0) {
    const int slots = occupancy_slots<T, N>();
    DGFS_HIP_TRY(hipMalloc(&h->Z, static_cast<size_t>(slots) * n * sizeof(cpx<T>)));
    DGFS_HIP_TRY(hipMalloc(&h->S, static_cast<size_t>(slots) * n * sizeof(T)));
    DGFS_HIP_TRY(hipMalloc(&h->W, static_cast<size_t>(slots) * 3 * N * N * sizeof(cpx<T>)));
    h->cap_slots = slots;
  }
  if (static_cast<size_t>(pts) > h->cap_pts) {
    (void)hipFree(h->F); (void)hipFree(h->fc); (void)hipFree(h->Qc);
    h->F = nullptr; h->fc = h->Qc = nullptr;
    DGFS_HIP_TRY(hipMalloc(&h->F, static_cast<size_t>(pts) * n * sizeof(cpx<T>)));
    DGFS_HIP_TRY(hipMalloc(&h->fc, static_cast<size_t>(pts) * n * sizeof(double)));
    DGFS_HIP_TRY(hipMalloc(&h->Qc, static_cast<size_t>(pts) * n * sizeof(double)));
    h->cap_pts = pts;
  }
  return DGFS_OK;
}

// points per launch: F + G^ (worst case one partial per radial node) + f + Q,
// plus the register-resident and slab kernels' spectra, Nyquist planes and T
// kept under 48 GiB of the 288 GB of HBM; launches are balanced in size
static int chunk_points(const VhsHandle* h, int total) {
  double per = static_cast<double>(h->n) * (2.0 * h->tsize * (1 + h->Nrho) + 16.0);
  if (use_rr(h)) per += 3.0 * h->N * h->N * h->Nrho * h->M * 2.0 * h->tsize + 2.0 * h->tsize * h->n;
  if (rr_kind(h) == 6) per += 2.0 * h->tsize * h->n * h->Nrho;  // the slab kernel's T
  const double budget = 48.0 * 1024.0 * 1024.0 * 1024.0;
  const int cap = std::max(1, static_cast<int>(budget / per));
  if (total <= cap) return std::max(total, 1);
  const int nchunks = (total + cap - 1) / cap;
  return (total + nchunks - 1) / nchunks;
}

template <typename T, int N>
static int eval_impl(VhsHandle* h, const double* f, double* Q, bool aosoa, int nupts, int ldim,
                     int vsize, int e0, int ne, int npts, hipStream_t st) {
  constexpr int n = N * N * N;
  if (aosoa) {
    const int total = nupts * ne;
    if (total == 0) return DGFS_OK;
    const int per_chunk_elems =
        std::max(1, chunk_points(h, total) / std::max(nupts, 1));
    for (int c0 = 0; c0 < ne; c0 += per_chunk_elems) {
      const int cne = std::min(per_chunk_elems, ne - c0);
      const int pts = cne * nupts;
      int rc = ensure_ws<T, N>(h, pts);
      if (rc) return rc;
      rc = launch_aosoa_gather(f, h->fc, nupts, ldim, vsize, e0 + c0, cne, st);
      if (rc) return rc;
      rc = run_chunk<T, N>(h, h->fc, h->Qc, pts, st);
      if (rc) return rc;
      rc = launch_aosoa_scatter(h->Qc, Q, nupts, ldim, vsize, e0 + c0, cne, st);
      if (rc) return rc;
    }
    return DGFS_OK;
  }
  const int chunk = chunk_points(h, npts);
  for (int p0 = 0; p0 < npts; p0 += chunk) {
    const int pts = std::min(chunk, npts - p0);
    int rc = ensure_ws<T, N>(h, pts);
    if (rc) return rc;
    rc = run_chunk<T, N>(h, f + static_cast<size_t>(p0) * n, Q + static_cast<size_t>(p0) * n, pts, st);
    if (rc) return rc;
  }
  return DGFS_OK;
}

template <typename T>
static int dispatch(VhsHandle* h, const double* f, double* Q, bool aosoa, int nupts, int ldim,
                    int vsize, int e0, int ne, int npts, hipStream_t st) {
  switch (h->N) {
#define DGFS_CASE(NV) \
  case NV:            \
    return eval_impl<T, NV>(h, f, Q, aosoa, nupts, ldim, vsize, e0, ne, npts, st);
    DGFS_CASE(4) DGFS_CASE(6) DGFS_CASE(8) DGFS_CASE(12) DGFS_CASE(16) DGFS_CASE(24) DGFS_CASE(32)
#undef DGFS_CASE
    default:
      return fail(DGFS_ERR_UNSUPPORTED, "Nv=%d is not built", h->N);
  }
}

// two handles (velocity grids NA > NB, same precision) in one persistent gain
// launch; both take the multi-pass gain
// ---------------------------------------------------------------------------
// allocation-free evaluation
// ---------------------------------------------------------------------------
// Callers that cannot allocate device memory (code under stream capture,
// serving loops, harnesses that forbid hipMalloc) evaluate Q through
// dgfs_vhs_eval_contig_noalloc, whose workspace is carved out of this
// file-scope array.  Each GPU the library runs on holds its own copy, at its
// own address, until the process ends; host code never passes the array's
// name to a kernel, only the device address that hipGetSymbolAddress returns
// for the current device.
constexpr size_t kScratchBytes = static_cast<size_t>(256) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// address of the current device's scratch array, resolved once per device
static int resolve_scratch(void** out) {
  *out = nullptr;
  int dev = 0;
  DGFS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(DGFS_ERR_UNSUPPORTED, "device ordinal %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    DGFS_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(DGFS_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DGFS_OK;
}

constexpr size_t kScratchAlign = 256;
static size_t align_up(size_t b) { return (b + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

// one chunk's carve-out of the scratch array
struct ScratchPlan {
  int pts = 0, groups = 0, slots = 0;
  size_t items = 0;  // gain partial buffers the Ghat share holds
  size_t z = 0, s = 0, w = 0, F = 0, Ghat = 0, Ft = 0, Wny = 0, Tb = 0;
  size_t total() const { return z + s + w + F + Ghat + Ft + Wny + Tb; }
};

template <typename T, int N>
static ScratchPlan plan_scratch(const VhsHandle* h, int pts, int slots) {
  constexpr size_t n = static_cast<size_t>(N) * N * N;
  const bool rr = use_rr(h);
  int gain_slots = rr ? cu_count() : slots;
  if (!rr && h->slots_override > 0) gain_slots = std::min(gain_slots, h->slots_override);
  ScratchPlan p;
  p.pts = pts;
  p.slots = slots;
  p.groups = pick_groups(h, pts, gain_slots);
  if constexpr (slab_available<N>()) {
    if (rr_kind(h) == 6) {
      p.groups = pick_groups(h, pts * SlabShape<N>::NS, gain_slots, kSlabMinGroups);
      p.Tb = align_up(static_cast<size_t>(pts) * h->Nrho * n * sizeof(cpx<T>));
    }
  }
  p.z = align_up(static_cast<size_t>(slots) * n * sizeof(cpx<T>));
  if (!rr) {  // the multi-pass gain's per-slot field and Nyquist planes
    p.s = align_up(static_cast<size_t>(slots) * n * sizeof(T));
    p.w = align_up(static_cast<size_t>(slots) * 3 * N * N * sizeof(cpx<T>));
  }
  p.F = align_up(static_cast<size_t>(pts) * n * sizeof(cpx<T>));
  p.items = static_cast<size_t>(pts) * p.groups;
  p.Ghat = align_up(p.items * n * sizeof(cpx<T>));
  if (rr) {
    p.Ft = p.F;
    if (rr_kind(h) != 3)
      p.Wny = align_up(static_cast<size_t>(pts) * h->Nrho * h->M * 3 * N * N * sizeof(cpx<T>));
  }
  return p;
}

// the handle's own workspace, set aside while a call runs on the scratch
// array and put back however the call ends
struct NoAllocScope {
  VhsHandle* h;
  void *F, *Ghat, *Z, *S, *W, *Wny, *Ft, *Tb;
  double *fc, *Qc;
  size_t cap_pts, cap_items, cap_ft, cap_wny, cap_tb;
  int cap_slots, timing;
  explicit NoAllocScope(VhsHandle* hh)
      : h(hh), F(hh->F), Ghat(hh->Ghat), Z(hh->Z), S(hh->S), W(hh->W), Wny(hh->Wny), Ft(hh->Ft), Tb(hh->Tb),
        fc(hh->fc), Qc(hh->Qc), cap_pts(hh->cap_pts), cap_items(hh->cap_items), cap_ft(hh->cap_ft),
        cap_wny(hh->cap_wny), cap_tb(hh->cap_tb), cap_slots(hh->cap_slots), timing(hh->timing) {
    h->no_alloc = 1;
    h->timing = 0;  // the event pool may grow; no timing in this mode
  }
  void use(char* base, const ScratchPlan& p) {
    char* at = base;
    auto take = [&](size_t b) -> void* {
      void* r = b ? static_cast<void*>(at) : nullptr;
      at += b;
      return r;
    };
    h->Z = take(p.z);
    h->S = take(p.s);
    h->W = take(p.w);
    h->F = take(p.F);
    h->Ghat = take(p.Ghat);
    h->Ft = take(p.Ft);
    h->Wny = take(p.Wny);
    h->Tb = take(p.Tb);
    h->fc = h->Qc = nullptr;  // the contiguous form reads f and writes Q in place
    h->cap_slots = p.slots;
    h->cap_pts = static_cast<size_t>(p.pts);
    h->cap_items = p.items;
    h->cap_ft = p.Ft;
    h->cap_wny = p.Wny;
    h->cap_tb = p.Tb;
  }
  ~NoAllocScope() {
    h->F = F; h->Ghat = Ghat; h->Z = Z; h->S = S; h->W = W; h->Wny = Wny; h->Ft = Ft; h->Tb = Tb;
    h->fc = fc; h->Qc = Qc;
    h->cap_pts = cap_pts; h->cap_items = cap_items; h->cap_ft = cap_ft; h->cap_wny = cap_wny; h->cap_tb = cap_tb;
    h->cap_slots = cap_slots;
    h->timing = timing;
    h->no_alloc = 0;
  }
};

template <typename T, int N>
static int eval_noalloc_impl(VhsHandle* h, const double* f, double* Q, int npts, hipStream_t st) {
  constexpr size_t n = static_cast<size_t>(N) * N * N;
  if (!h->scratch)
    return fail(DGFS_ERR_NOT_INITIALIZED, "no scratch array on this device: %s", h->scratch_err.c_str());
  // per-slot buffers of the prep/finalize (and multi-pass gain) grids take at
  // most half of the array; the rest holds the chunk's spectra and partials
  int slots = std::min(occupancy_slots<T, N>(), cu_count());
  const size_t per_slot = n * (sizeof(cpx<T>) + sizeof(T)) + 3 * N * N * sizeof(cpx<T>) + 3 * kScratchAlign;
  slots = static_cast<int>(std::max<size_t>(1, std::min<size_t>(slots, kScratchBytes / 2 / per_slot)));
  // largest balanced chunk that fits; the last chunk may be smaller and
  // then split its points into more radial groups, so the gain partials'
  // share is sized for whichever of the two needs more
  auto plan_for = [&](int cap, int* chunk_out) {
    const int nchunks = (npts + cap - 1) / cap;
    const int chunk = (npts + nchunks - 1) / nchunks;
    const int last = npts - ((npts + chunk - 1) / chunk - 1) * chunk;
    ScratchPlan a = plan_scratch<T, N>(h, chunk, slots);
    const ScratchPlan b = plan_scratch<T, N>(h, last, slots);
    if (b.Ghat > a.Ghat) {
      a.Ghat = b.Ghat;
      a.items = b.items;
    }
    *chunk_out = chunk;
    return a;
  };
  int cap = npts, chunk = 0;
  ScratchPlan plan;
  for (;;) {
    if (cap < 1)
      return fail(DGFS_ERR_UNSUPPORTED,
                  "one point at Nv=%d (precision %d) needs %zu bytes of workspace, beyond the %zu-byte scratch array",
                  N, h->prec, plan_scratch<T, N>(h, 1, slots).total(), kScratchBytes);
    plan = plan_for(cap, &chunk);
    if (plan.total() <= kScratchBytes) break;
    const double over = static_cast<double>(plan.total()) / static_cast<double>(kScratchBytes);
    cap = std::min(cap - 1, static_cast<int>(cap / over));
  }
  NoAllocScope scope(h);
  scope.use(static_cast<char*>(h->scratch), plan);
  for (int p0 = 0; p0 < npts; p0 += chunk) {
    const int pts = std::min(chunk, npts - p0);
    const int rc = run_chunk<T, N>(h, f + static_cast<size_t>(p0) * n, Q + static_cast<size_t>(p0) * n, pts, st);
    if (rc) return rc;
  }
  return DGFS_OK;
}

template <typename T>
static int dispatch_noalloc(VhsHandle* h, const double* f, double* Q, int npts, hipStream_t st) {
  switch (h->N) {
#define DGFS_CASE(NV) \
  case NV:            \
    return eval_noalloc_impl<T, NV>(h, f, Q, npts, st);
    DGFS_CASE(4) DGFS_CASE(6) DGFS_CASE(8) DGFS_CASE(12) DGFS_CASE(16) DGFS_CASE(24) DGFS_CASE(32)
#undef DGFS_CASE
    default:
      return fail(DGFS_ERR_UNSUPPORTED, "Nv=%d is not built", h->N);
  }
}

template <typename T, int NA, int NB>
static int run_pair(VhsHandle* ha, const double* fa, double* Qa, int na, VhsHandle* hb, const double* fb, double* Qb,
                    int nb, hipStream_t st) {
  int rc = ensure_ws<T, NA>(ha, na);
  if (!rc) rc = ensure_ws<T, NB>(hb, nb);
  if (rc) return rc;
  const int slots = std::min(ha->cap_slots, hb->cap_slots);
  const int gA = pick_groups(ha, na + nb, slots), gB = pick_groups(hb, na + nb, slots);
  GainArgs<T> ga, gb;
  rc = stage_prep<T, NA>(ha, fa, na, gA, slots, false, &ga, st);
  if (!rc) rc = stage_prep<T, NB>(hb, fb, nb, gB, slots, false, &gb, st);
  if (rc) return rc;
  ha->last_algorithm = hb->last_algorithm = 1;
  {
    KTimerT<VhsHandle> t1(ha, 1, st);
    hipLaunchKernelGGL((k_vhs_gain_pair<T, NA, NB>), dim3(std::min(ga.nitems + gb.nitems, slots)),
                       dim3(Shape<NA>::NT), 0, st, ga, gb);
    DGFS_LAUNCH_CHECK("vhs_gain_pair");
    t1.stop();
  }
  rc = stage_finish<T, NA>(ha, fa, Qa, na, gA, st);
  if (!rc) rc = stage_finish<T, NB>(hb, fb, Qb, nb, gB, st);
  return rc;
}

// Two register-resident grids: the two pipelines (prep, Nyquist planes, gain,
// finalize, projection) run back to back on the caller's stream, each gain
// grid filling the chip.  The two kernels need different workgroup sizes (512
// threads at 32^3, 576 at 24^3), so they cannot share one launch; run side by
// side on a CU split over a second stream they measured 13% slower than back
// to back on MI355X (90.4 against 79.9 ms for 4096 points at 32^3 plus 1024
// at 24^3), because both are bound by the CU's VALU and LDS, not by launch
// gaps.  Q is bitwise that of two separate calls.
static int eval_pair_rr(VhsHandle* ha, const double* fa, double* Qa, int na, VhsHandle* hb, const double* fb,
                        double* Qb, int nb, hipStream_t st) {
  int rc = (ha->prec == 64) ? dispatch<double>(ha, fa, Qa, false, 0, 0, 0, 0, 0, na, st)
                            : dispatch<float>(ha, fa, Qa, false, 0, 0, 0, 0, 0, na, st);
  if (rc) return rc;
  rc = (hb->prec == 64) ? dispatch<double>(hb, fb, Qb, false, 0, 0, 0, 0, 0, nb, st)
                        : dispatch<float>(hb, fb, Qb, false, 0, 0, 0, 0, 0, nb, st);
  if (rc) return rc;
  ha->last_algorithm = rr_kind(ha);
  hb->last_algorithm = rr_kind(hb);
  return DGFS_OK;
}

template <typename T>
static int dispatch_pair(VhsHandle* ha, const double* fa, double* Qa, int na, VhsHandle* hb, const double* fb,
                         double* Qb, int nb, hipStream_t st) {
  if (use_rr(ha) && use_rr(hb)) return eval_pair_rr(ha, fa, Qa, na, hb, fb, Qb, nb, st);
  const int key = ha->N * 100 + hb->N;
  switch (key) {
    case 3224: return run_pair<T, 32, 24>(ha, fa, Qa, na, hb, fb, Qb, nb, st);
    case 3216: return run_pair<T, 32, 16>(ha, fa, Qa, na, hb, fb, Qb, nb, st);
    case 2416: return run_pair<T, 24, 16>(ha, fa, Qa, na, hb, fb, Qb, nb, st);
    default: return fail(DGFS_ERR_UNSUPPORTED, "paired grids %d and %d are not built (32+24, 32+16, 24+16)",
                         ha->N, hb->N);
  }
}

static VhsHandle* as_handle(void* p) {
  auto* h = static_cast<VhsHandle*>(p);
  if (!h || h->magic != 0x76687331) return nullptr;
  return h;
}

}  // namespace dgfs

using namespace dgfs;

extern "C" const char* dgfs_last_error(void) { return last_error_ref().c_str(); }
extern "C" const char* dgfs_version(void) { return "dgfs_hip 0.1 (gfx950)"; }

extern "C" int dgfs_vhs_create(const dgfs_vhs_desc* d, void** out) {
  if (!d || !out) return fail(DGFS_ERR_INVALID_VALUE, "null descriptor or handle pointer");
  *out = nullptr;
  if (!supported_N(d->Nv))
    return fail(DGFS_ERR_UNSUPPORTED, "Nv=%d unsupported (built: 4 6 8 12 16 24 32)", d->Nv);
  if (d->Nrho < 1 || d->M < 1 || !d->qz || !d->qw || !d->sz)
    return fail(DGFS_ERR_INVALID_VALUE, "Nrho, M must be >= 1 and qz/qw/sz non-null");
  if (d->precision != 64 && d->precision != 32)
    return fail(DGFS_ERR_INVALID_VALUE, "precision must be 64 or 32 (got %d)", d->precision);
  if (!(d->L > 0) || !(d->cw > 0)) return fail(DGFS_ERR_INVALID_VALUE, "L and cw must be positive");

  auto h = std::make_unique<VhsHandle>();
  h->N = d->Nv;
  h->n = d->Nv * d->Nv * d->Nv;
  h->Nrho = d->Nrho;
  h->M = d->M;
  h->prec = d->precision;
  h->project = d->project ? 1 : 0;
  h->L = d->L; h->gamma = d->gamma; h->prefac = d->prefac; h->sw = d->sw; h->cw = d->cw;
  h->tsize = (d->precision == 64) ? 8 : 4;
  const int N = h->N, n = h->n;

  // b1_p(l) = qz_p^(gamma+2) 4 pi sinc((pi/L)(qz_p/2)|l|),
  // b2(l) = sum_p qw_p qz_p^(gamma+2) 16 pi^2 sinc((pi/L) qz_p |l|)
  // with sinc(s) = sin(s + 1e-15)/(s + 1e-15)  (vhs-gll.mako:11-14, 71-101)
  std::vector<double> b1(static_cast<size_t>(h->Nrho) * n), b2(n, 0.0);
  const double pi = 3.14159265358979323846;
  auto sinc = [](double s) { return std::sin(s + 1e-15) / (s + 1e-15); };
  for (int r = 0; r < n; ++r) {
    const int I = r / (N * N), J = (r / N) % N, K = r % N;
    const double lx = (I < N / 2) ? I : I - N, ly = (J < N / 2) ? J : J - N, lz = (K < N / 2) ? K : K - N;
    const double ln = std::sqrt(lx * lx + ly * ly + lz * lz);
    double acc = 0.0;
    for (int p = 0; p < h->Nrho; ++p) {
      const double qz = d->qz[p], pw = std::pow(qz, h->gamma + 2.0);
      b1[static_cast<size_t>(p) * n + r] = pw * 4.0 * pi * sinc(pi / h->L * qz / 2.0 * ln);
      acc += d->qw[p] * pw * 16.0 * pi * pi * sinc(pi / h->L * qz * ln);
    }
    b2[r] = acc;
  }
  std::vector<double> alpha(h->Nrho), wgt(h->Nrho);
  for (int p = 0; p < h->Nrho; ++p) {
    alpha[p] = pi / h->L * d->qz[p] / 2.0;
    wgt[p] = 2.0 * d->qw[p] * h->sw / static_cast<double>(n);
  }
  auto upload = [&](void** dst, const std::vector<double>& src) -> int {
    if (h->prec == 64) {
      DGFS_HIP_TRY(hipMalloc(dst, src.size() * sizeof(double)));
      DGFS_HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
      std::vector<float> tmp(src.begin(), src.end());
      DGFS_HIP_TRY(hipMalloc(dst, tmp.size() * sizeof(float)));
      DGFS_HIP_TRY(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return DGFS_OK;
  };
  int rc = upload(&h->b1, b1);
  if (!rc) rc = upload(&h->b2, b2);
  if (rc) return rc;
  DGFS_HIP_TRY(hipMalloc(&h->alpha, h->Nrho * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&h->wgt, h->Nrho * sizeof(double)));
  DGFS_HIP_TRY(hipMalloc(&h->szd, h->M * 3 * sizeof(double)));
  DGFS_HIP_TRY(hipMemcpy(h->alpha, alpha.data(), h->Nrho * sizeof(double), hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(h->wgt, wgt.data(), h->Nrho * sizeof(double), hipMemcpyHostToDevice));
  DGFS_HIP_TRY(hipMemcpy(h->szd, d->sz, h->M * 3 * sizeof(double), hipMemcpyHostToDevice));
  {
    // separable phase factors of every (p, q), evaluated in FP64 on the host
    std::vector<double> tab(static_cast<size_t>(h->Nrho) * h->M * 3 * N * 2);
    for (int p = 0; p < h->Nrho; ++p)
      for (int q = 0; q < h->M; ++q)
        for (int dd = 0; dd < 3; ++dd)
          for (int k = 0; k < N; ++k) {
            const double l = (k < N / 2) ? k : k - N;
            const double ang = alpha[p] * d->sz[q * 3 + dd] * l;
            const size_t o = ((static_cast<size_t>(p) * h->M + q) * 3 * N + dd * N + k) * 2;
            tab[o] = std::cos(ang);
            tab[o + 1] = std::sin(ang);
          }
    rc = upload(&h->ephtab, tab);
    if (rc) return rc;
  }

  // (C C^T)^{-1} / cw^2 for the projection: A_jk = sum_r phi_j phi_k cw
  double A[25] = {};
  const double hh = 2.0 * h->L / N, c0 = -h->L + 0.5 * hh;
  for (int r = 0; r < n; ++r) {
    const int I = r / (N * N), J = (r / N) % N, K = r % N;
    const double cx = c0 + I * hh, cy = c0 + J * hh, cz = c0 + K * hh;
    const double phi[5] = {1.0, cx, cy, cz, cx * cx + cy * cy + cz * cz};
    for (int j = 0; j < 5; ++j)
      for (int k = 0; k < 5; ++k) A[j * 5 + k] += phi[j] * phi[k] * h->cw * h->cw;
  }
  if (invert5(A, h->Ainv)) return fail(DGFS_ERR_INVALID_VALUE, "singular invariant Gram matrix");
  // the allocation-free entry point's scratch array on this device; a
  // failure here only disables that entry point
  if (resolve_scratch(&h->scratch) != DGFS_OK) {
    h->scratch = nullptr;
    h->scratch_err = last_error_ref();
    last_error_ref().clear();
  }
  *out = h.release();
  return DGFS_OK;
}

extern "C" int dgfs_vhs_destroy(void* p) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  resolve_timing(h);
  for (hipEvent_t e : h->pool) (void)hipEventDestroy(e);
  h->pool.clear();
  free_ws(h);
  (void)hipFree(h->b1); (void)hipFree(h->b2); (void)hipFree(h->alpha); (void)hipFree(h->wgt);
  (void)hipFree(h->szd);
  (void)hipFree(h->ephtab);
  h->magic = 0;
  delete h;
  return DGFS_OK;
}

extern "C" int dgfs_vhs_eval(void* p, const double* f, double* Q, int nupts, int ldim, int vsize,
                             int neles, int e0, int ne, void* stream) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (!f || !Q) return fail(DGFS_ERR_INVALID_VALUE, "null distribution pointer");
  if (f == Q) return fail(DGFS_ERR_INVALID_VALUE, "f and Q must not alias");
  if (vsize != h->n) return fail(DGFS_ERR_INVALID_VALUE, "vsize %d != Nv^3 = %d", vsize, h->n);
  if (nupts < 0 || neles < 0 || e0 < 0 || ne < 0 || e0 + ne > neles)
    return fail(DGFS_ERR_INVALID_VALUE, "element range [%d, %d) outside [0, %d)", e0, e0 + ne, neles);
  const long long nparr = ((static_cast<long long>(neles) + 31) / 32) * 32;
  if (static_cast<long long>(ldim) < nparr * vsize)
    return fail(DGFS_ERR_INVALID_VALUE, "ldim %d smaller than the AoSoA row %lld", ldim, nparr * vsize);
  auto st = static_cast<hipStream_t>(stream);
  return (h->prec == 64) ? dispatch<double>(h, f, Q, true, nupts, ldim, vsize, e0, ne, 0, st)
                         : dispatch<float>(h, f, Q, true, nupts, ldim, vsize, e0, ne, 0, st);
}

extern "C" int dgfs_vhs_eval_contig(void* p, const double* f, double* Q, int npts, void* stream) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (!f || !Q) return fail(DGFS_ERR_INVALID_VALUE, "null distribution pointer");
  if (f == Q) return fail(DGFS_ERR_INVALID_VALUE, "f and Q must not alias");
  if (npts < 0) return fail(DGFS_ERR_INVALID_VALUE, "npts < 0");
  if (npts == 0) return DGFS_OK;
  auto st = static_cast<hipStream_t>(stream);
  return (h->prec == 64) ? dispatch<double>(h, f, Q, false, 0, 0, 0, 0, 0, npts, st)
                         : dispatch<float>(h, f, Q, false, 0, 0, 0, 0, 0, npts, st);
}

extern "C" int dgfs_vhs_eval_contig_noalloc(void* p, const double* f, double* Q, int npts, void* stream) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (!f || !Q) return fail(DGFS_ERR_INVALID_VALUE, "null distribution pointer");
  if (f == Q) return fail(DGFS_ERR_INVALID_VALUE, "f and Q must not alias");
  if (npts < 0) return fail(DGFS_ERR_INVALID_VALUE, "npts < 0");
  if (npts == 0) return DGFS_OK;
  if (h->no_alloc) return fail(DGFS_ERR_INVALID_VALUE, "handle already inside an allocation-free call");
  auto st = static_cast<hipStream_t>(stream);
  return (h->prec == 64) ? dispatch_noalloc<double>(h, f, Q, npts, st) : dispatch_noalloc<float>(h, f, Q, npts, st);
}

extern "C" int dgfs_vhs_get_dims(void* p, int* Nv, int* Nrho, int* M, int* precision, double* L) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (Nv) *Nv = h->N;
  if (Nrho) *Nrho = h->Nrho;
  if (M) *M = h->M;
  if (precision) *precision = h->prec;
  if (L) *L = h->L;
  return DGFS_OK;
}

extern "C" int dgfs_vhs_get_config(void* p, int* ngroups, int* nslots, int* threads) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (ngroups) *ngroups = h->last_groups;
  if (nslots) *nslots = h->last_slots;
  if (threads) *threads = h->threads;
  return DGFS_OK;
}

extern "C" int dgfs_vhs_set_groups(void* p, int ngroups) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (ngroups < 0) return fail(DGFS_ERR_INVALID_VALUE, "ngroups < 0");
  h->groups_override = ngroups;
  return DGFS_OK;
}

extern "C" int dgfs_vhs_set_slots(void* p, int nslots) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (nslots < 0) return fail(DGFS_ERR_INVALID_VALUE, "nslots < 0");
  h->slots_override = nslots;
  return DGFS_OK;
}

extern "C" int dgfs_vhs_set_timing(void* p, int on) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  resolve_timing(h);
  h->timing = on ? 1 : 0;
  for (int k = 0; k < 5; ++k) {
    h->acc_ms[k] = 0.0;
    h->acc_n[k] = 0;
  }
  return DGFS_OK;
}

extern "C" int dgfs_vhs_get_timing(void* p, double* ms, int* counts) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (!ms || !counts) return fail(DGFS_ERR_INVALID_VALUE, "null output arrays");
  resolve_timing(h);
  for (int k = 0; k < 5; ++k) {
    ms[k] = h->acc_ms[k];
    counts[k] = h->acc_n[k];
  }
  return DGFS_OK;
}

extern "C" int dgfs_vhs_set_algorithm(void* p, int algorithm) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (algorithm < 0 || algorithm > 6) return fail(DGFS_ERR_INVALID_VALUE, "algorithm must be 0 to 6");
  if ((algorithm == 2 || algorithm == 3 || algorithm == 5) && !rr2_ok(h))
    return fail(DGFS_ERR_UNSUPPORTED, "register-resident kernel (two pencils per thread): FP32 and Nv in {16, 32} only");
  if (algorithm == 4 && !rr1_ok(h))
    return fail(DGFS_ERR_UNSUPPORTED, "register-resident kernel (one pencil per thread): FP32 and Nv in {16, 24} only");
  if (algorithm == 6 && !slab_ok(h))
    return fail(DGFS_ERR_UNSUPPORTED, "slab kernel: Nv in {16, 24, 32} only");
  h->algorithm = algorithm;
  return DGFS_OK;
}

extern "C" int dgfs_vhs_get_algorithm(void* p, int* algorithm) {
  VhsHandle* h = as_handle(p);
  if (!h) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (!algorithm) return fail(DGFS_ERR_INVALID_VALUE, "null output");
  *algorithm = h->last_algorithm;
  return DGFS_OK;
}

// Two independent batches on two velocity grids (for instance 32^3 and 24^3
// points of a mixed run) in one persistent gain launch.  Batches larger than
// one workspace chunk are evaluated as separate launches.
extern "C" int dgfs_vhs_eval_pair(void* pa, const double* fa, double* Qa, int na, void* pb, const double* fb,
                                  double* Qb, int nb, void* stream) {
  VhsHandle* ha = as_handle(pa);
  VhsHandle* hb = as_handle(pb);
  if (!ha || !hb) return fail(DGFS_ERR_NOT_INITIALIZED, "invalid VHS handle");
  if (ha == hb) return fail(DGFS_ERR_INVALID_VALUE, "the two batches need two handles");
  if (!fa || !Qa || !fb || !Qb || na < 0 || nb < 0) return fail(DGFS_ERR_INVALID_VALUE, "bad batch arguments");
  if (fa == Qa || fb == Qb) return fail(DGFS_ERR_INVALID_VALUE, "f and Q may not alias");
  if (ha->prec != hb->prec) return fail(DGFS_ERR_INVALID_VALUE, "paired handles need the same precision");
  auto st = static_cast<hipStream_t>(stream);
  if (na == 0 || nb == 0 || na > chunk_points(ha, na) || nb > chunk_points(hb, nb)) {
    int rc = dgfs_vhs_eval_contig(pa, fa, Qa, na, stream);
    return rc ? rc : dgfs_vhs_eval_contig(pb, fb, Qb, nb, stream);
  }
  if (ha->N < hb->N) {
    std::swap(ha, hb);
    std::swap(fa, fb);
    std::swap(Qa, Qb);
    std::swap(na, nb);
  }
  return (ha->prec == 64) ? dispatch_pair<double>(ha, fa, Qa, na, hb, fb, Qb, nb, st)
                          : dispatch_pair<float>(ha, fa, Qa, na, hb, fb, Qb, nb, st);
}
