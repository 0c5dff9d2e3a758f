// Complex arithmetic and thread-local mixed-radix FFTs for gfx950.
//
// A pencil of N complex values lives in one thread's registers; N is a
// compile-time constant (N = 2^a 3^b) so every butterfly and twiddle is
// unrolled into straight-line code.  Twiddles are evaluated at compile time
// by a constexpr sine/cosine (argument reduction to [0, pi/4] by the octant
// symmetries of the circle, then a Taylor series to 1 ulp), so no table loads
// sit in the inner loops.
#pragma once
#include <hip/hip_runtime.h>

namespace dgfs {

// naturally aligned (2 * sizeof(T)) so that LDS and global accesses of one
// element are single 64-bit (FP32) or 128-bit (FP64) operations with wide
// immediate offsets
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b) { return cpx<T>{a, b}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return {s * a.x, s * a.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conj(cpx<T> a) { return {a.x, -a.y}; }
// multiply by +i
template <typename T>
__host__ __device__ __forceinline__ cpx<T> mul_i(cpx<T> a) { return {-a.y, a.x}; }

// ---------------------------------------------------------------------------
// compile-time trigonometry of rational angles 2*pi*k/n
// ---------------------------------------------------------------------------
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
// cos and sin of 2*pi*k/n.  The angle is split on the integer pair into an
// octant o = floor(8k/n) and a remainder phi in [0, pi/4); odd octants use
// the complement pi/4 - phi so the Taylor argument never exceeds pi/4, and
// the quadrant o/2 is applied by exact sign/swap rotations.
constexpr void cs_rational(long k, long n, double& c, double& s) {
  k %= n;
  if (k < 0) k += n;
  const long e = 8 * k;
  const int oct = static_cast<int>(e / n);
  const long rem = e - oct * n;
  double cq = 0.0, sq = 0.0;  // cos/sin of the angle within the quadrant
  if (oct & 1) {
    const double comp = (kPi / 4.0) * static_cast<double>(n - rem) / static_cast<double>(n);
    cq = taylor_sin(comp);
    sq = taylor_cos(comp);
  } else {
    const double phi = (kPi / 4.0) * static_cast<double>(rem) / static_cast<double>(n);
    cq = taylor_cos(phi);
    sq = taylor_sin(phi);
  }
  switch ((oct / 2) & 3) {
    case 0: c = cq; s = sq; break;
    case 1: c = -sq; s = cq; break;
    case 2: c = -cq; s = -sq; break;
    default: c = sq; s = -cq; break;
  }
}
constexpr double cos_rational(long k, long n) { double c = 0, s = 0; cs_rational(k, n, c, s); return c; }
constexpr double sin_rational(long k, long n) { double c = 0, s = 0; cs_rational(k, n, c, s); return s; }

template <typename T, int SIGN>
__device__ __forceinline__ cpx<T> twmul(cpx<T> a, T c, T s) {
  // a * (c + i s)
  return {a.x * c - a.y * s, a.x * s + a.y * c};
}

// ---------------------------------------------------------------------------
// radix of the current decimation step: 4 where possible, then 2, then 3
// ---------------------------------------------------------------------------
template <int N>
struct first_radix {
  static constexpr int value = (N % 4 == 0 && N > 4) ? 4 : (N % 2 == 0) ? 2 : (N % 3 == 0) ? 3 : N;
};

// in-register DFT of length N (stride-1 in and out); SIGN = -1 forward,
// +1 inverse, both unnormalised.  Decimation in time: x[r + R*m] gives R
// sub-DFTs of length M = N/R, recombined by R-point butterflies with
// twiddles W_N^{r k} that are compile-time constants.  TAN: radix-4
// butterflies in the tangent form (twpm), 6.4 % fewer operations at N = 32;
// measured on MI355X it speeds the register-resident gain kernels (1.6 %)
// and slows the LDS-staged prep / finalize / Nyquist kernels (their FFTs
// then hold more literal constants), so only the gain kernels take it.
template <typename T, int N, int SIGN, bool TAN = false>
struct FFT {
  static_assert(N >= 1, "bad length");
  static constexpr int R = first_radix<N>::value;
  static constexpr int M = N / R;
  static_assert(N <= 4 || R * M == N, "N must be 2^a 3^b");

  template <int k, int r>
  __device__ __forceinline__ static void twiddle_in(const cpx<T>* y, cpx<T>* v) {
    if constexpr (r < R) {
      if constexpr (r == 0 || k == 0) {
        v[r] = y[r * M + k];
      } else {
        constexpr T c = static_cast<T>(cos_rational(static_cast<long>(r) * k, N));
        constexpr T s = static_cast<T>(SIGN * sin_rational(static_cast<long>(r) * k, N));
        // the general product even for quarter turns: written as swaps and
        // sign changes they made the FP64 two-species multi-pass gain 39 %
        // slower on MI355X (302 -> 419 ms per 2048 points at 24^3), with no
        // gain elsewhere
        v[r] = twmul<T, SIGN>(y[r * M + k], c, s);
      }
      twiddle_in<k, r + 1>(y, v);
    }
  }

  __device__ __forceinline__ static T fmadd(T a, T b, T c) {
    if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c); else return __builtin_fmaf(a, b, c);
  }

  // a * W^e, W = exp(SIGN 2 pi i / N)
  template <int e>
  __device__ __forceinline__ static cpx<T> tw(cpx<T> a) {
    constexpr T c = static_cast<T>(cos_rational(e, N));
    constexpr T s = static_cast<T>(SIGN * sin_rational(e, N));
    if constexpr (e % N == 0) {
      return a;
    } else if constexpr (c == T(0) && s == T(1)) {
      return cpx<T>{-a.y, a.x};
    } else if constexpr (c == T(0) && s == T(-1)) {
      return cpx<T>{a.y, -a.x};
    } else if constexpr (s == T(0) && c == T(-1)) {
      return cpx<T>{-a.x, -a.y};
    } else {
      return twmul<T, SIGN>(a, c, s);
    }
  }

  // sum = p + W^e a, dif = p - W^e a.  With W^e = c + i s and |c| >= |s|,
  // W^e a = c (a.x - t a.y, a.y + t a.x), t = s / c; otherwise
  // W^e a = s (t a.x - a.y, t a.y + a.x), t = c / s.  The scaled vector u
  // costs two fused multiply-adds and each output two more: six operations
  // where the product and the two sums take eight (the tangent form of a
  // twiddled butterfly; t is a compile-time constant).
  template <int e>
  __device__ __forceinline__ static void twpm(cpx<T> p, cpx<T> a, cpx<T>& sum, cpx<T>& dif) {
    constexpr double cd = cos_rational(e, N), sd = SIGN * sin_rational(e, N);
    constexpr T c = static_cast<T>(cd), s = static_cast<T>(sd);
    if constexpr (e % N == 0 || (c == T(0) && (s == T(1) || s == T(-1))) || (s == T(0) && c == T(-1))) {
      const cpx<T> b = tw<e>(a);
      sum = p + b;
      dif = p - b;
    } else if constexpr ((cd < 0 ? -cd : cd) >= (sd < 0 ? -sd : sd)) {
      constexpr T t = static_cast<T>(sd / cd);
      const cpx<T> u{fmadd(-t, a.y, a.x), fmadd(t, a.x, a.y)};
      sum = cpx<T>{fmadd(c, u.x, p.x), fmadd(c, u.y, p.y)};
      dif = cpx<T>{fmadd(-c, u.x, p.x), fmadd(-c, u.y, p.y)};
    } else {
      constexpr T t = static_cast<T>(cd / sd);
      const cpx<T> u{fmadd(t, a.x, -a.y), fmadd(t, a.y, a.x)};
      sum = cpx<T>{fmadd(s, u.x, p.x), fmadd(s, u.y, p.y)};
      dif = cpx<T>{fmadd(-s, u.x, p.x), fmadd(-s, u.y, p.y)};
    }
  }

  // radix-4 butterfly k with the twiddles folded into the sums:
  // t0,1 = a0 +- w2 a2; b1 = w1 a1; t2,3 = b1 +- w3 a3; X = t0 +- t2,
  // t1 +- (SIGN i) t3 (w_r = W^{r k})
  template <int k>
  __device__ __forceinline__ static void radix4(cpx<T>* x, const cpx<T>* y) {
    cpx<T> t0, t1, t2, t3;
    twpm<2 * k>(y[k], y[2 * M + k], t0, t1);
    twpm<3 * k>(tw<k>(y[M + k]), y[3 * M + k], t2, t3);
    const cpx<T> t3r = (SIGN > 0) ? mk<T>(-t3.y, t3.x) : mk<T>(t3.y, -t3.x);
    x[k] = t0 + t2;
    x[k + 2 * M] = t0 - t2;
    x[k + M] = t1 + t3r;
    x[k + 3 * M] = t1 - t3r;
  }

  template <int k>
  __device__ __forceinline__ static void butterflies(cpx<T>* x, const cpx<T>* y) {
    if constexpr (k < M) {
      if constexpr (R == 4 && TAN) {
        radix4<k>(x, y);
      } else {
        cpx<T> v[R];
        twiddle_in<k, 0>(y, v);
        FFT<T, R, SIGN, TAN>::run(v);
#pragma unroll
        for (int s = 0; s < R; ++s) x[k + M * s] = v[s];
      }
      butterflies<k + 1>(x, y);
    }
  }

  __device__ __forceinline__ static void run(cpx<T>* x) {
    if constexpr (N == 1) {
      return;
    } else if constexpr (N == 2) {
      cpx<T> a = x[0], b = x[1];
      x[0] = a + b;
      x[1] = a - b;
    } else if constexpr (N == 3) {
      constexpr T c = static_cast<T>(-0.5);
      constexpr T s = static_cast<T>(SIGN * 0.86602540378443864676372317075294);
      cpx<T> a = x[0], b = x[1], d = x[2];
      cpx<T> t = b + d, u = b - d;
      x[0] = a + t;
      cpx<T> m1 = a + c * t;
      cpx<T> m2 = mk<T>(-s * u.y, s * u.x);
      x[1] = m1 + m2;
      x[2] = m1 - m2;
    } else if constexpr (N == 4) {
      cpx<T> a = x[0], b = x[1], c = x[2], d = x[3];
      cpx<T> t0 = a + c, t1 = a - c, t2 = b + d, t3 = b - d;
      cpx<T> t3r = (SIGN > 0) ? mk<T>(-t3.y, t3.x) : mk<T>(t3.y, -t3.x);
      x[0] = t0 + t2;
      x[2] = t0 - t2;
      x[1] = t1 + t3r;
      x[3] = t1 - t3r;
    } else {
      cpx<T> y[N];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int m = 0; m < M; ++m) y[r * M + m] = x[r + R * m];
#pragma unroll
      for (int r = 0; r < R; ++r) FFT<T, M, SIGN, TAN>::run(y + r * M);
      butterflies<0>(x, y);
    }
  }
};

}  // namespace dgfs
