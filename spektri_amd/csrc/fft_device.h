// Device-side FFT building blocks for gfx950 (CDNA4), FP32 throughout.
//
// * cpx: interleaved complex f32 (8 bytes, matches Complex<f32> and cf32le).
// * Compile-time twiddles: exp(SIGN * 2 pi i E / N) evaluated in double at
//   compile time (Taylor series after exact quadrant reduction) and rounded
//   to f32 literals; trivial twiddles (1, -1, +-i) cost no multiplies.
// * dft<R, SIGN>: an in-register DFT of R = 2^m points, built recursively as
//   R = R1 x R2 decimation in time.  SIGN = -1 is the forward transform
//   (rustfft forward, e^{-2 pi i nk/N}), SIGN = +1 the unnormalised inverse.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace spk {

struct alignas(8) cpx {
  float x, y;
};

__device__ __forceinline__ cpx cadd(cpx a, cpx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cpx csub(cpx a, cpx b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cpx cmul(cpx a, cpx b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
__device__ __forceinline__ cpx cmulc(cpx a, cpx b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
__device__ __forceinline__ cpx cscale(cpx a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cpx cconj(cpx a) { return {a.x, -a.y}; }
// multiply by SIGN * i
template <int SIGN>
__device__ __forceinline__ cpx mul_si(cpx a) {
  if constexpr (SIGN > 0) return {-a.y, a.x};
  else return {a.y, -a.x};
}

// ---------------------------------------------------------------------------
// Compile-time trigonometry
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_taylor(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 16; ++n) {
    term *= -x2 / double((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double ct_cos_taylor(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 16; ++n) {
    term *= -x2 / double((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi e / n, quadrant reduced exactly in integers.
constexpr double ct_cos2pi(long e, long n) {
  e %= n;
  if (e < 0) e += n;
  long q = (4 * e) / n;           // quadrant 0..3
  long r = 4 * e - q * n;         // remainder in [0, n)
  double phi = (kPi / 2) * double(r) / double(n);
  double c = ct_cos_taylor(phi), s = ct_sin_taylor(phi);
  switch (q) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
  }
}
constexpr double ct_sin2pi(long e, long n) {
  e %= n;
  if (e < 0) e += n;
  long q = (4 * e) / n;
  long r = 4 * e - q * n;
  double phi = (kPi / 2) * double(r) / double(n);
  double c = ct_cos_taylor(phi), s = ct_sin_taylor(phi);
  switch (q) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
  }
}

// a * exp(SIGN * 2 pi i E / N), E and N compile-time.
template <int N, int E, int SIGN>
__device__ __forceinline__ cpx twc(cpx a) {
  constexpr int e = ((E % N) + N) % N;
  if constexpr (e == 0) {
    return a;
  } else if constexpr (4 * e == N) {
    return mul_si<SIGN>(a);
  } else if constexpr (2 * e == N) {
    return {-a.x, -a.y};
  } else if constexpr (4 * e == 3 * N) {
    return mul_si<-SIGN>(a);
  } else {
    constexpr float c = float(ct_cos2pi(e, N));
    constexpr float s = float(SIGN * ct_sin2pi(e, N));
    return {a.x * c - a.y * s, a.x * s + a.y * c};
  }
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// ---------------------------------------------------------------------------
// In-register DFT, natural order in and out.
// ---------------------------------------------------------------------------
// With CONJ_HI the outputs R/2 .. R-1 come out conjugated: the last
// butterflies form their imaginary parts as reversed differences, at no cost.
template <int R, int SIGN, bool CONJ_HI = false>
__device__ __forceinline__ void dft(cpx (&x)[R]) {
  static_assert((R & (R - 1)) == 0, "power of two radices only");
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    cpx a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = CONJ_HI ? cpx{a.x - b.x, b.y - a.y} : csub(a, b);
  } else if constexpr (R == 4) {
    cpx s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
    cpx s13 = cadd(x[1], x[3]), d13 = mul_si<SIGN>(csub(x[1], x[3]));
    x[0] = cadd(s02, s13);
    x[2] = CONJ_HI ? cpx{s02.x - s13.x, s13.y - s02.y} : csub(s02, s13);
    x[1] = cadd(d02, d13);
    x[3] = CONJ_HI ? cpx{d02.x - d13.x, d13.y - d02.y} : csub(d02, d13);
  } else {
    // n = R1*n2 + n1, k = k2 + R2*k1
    constexpr int R1 = (R >= 16) ? 4 : 2;
    constexpr int R2 = R / R1;
    cpx a[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) a[n1][n2] = x[R1 * n2 + n1];
      dft<R2, SIGN>(a[n1]);
    }
    static_for<1, R1>([&](auto n1c) {
      constexpr int n1 = decltype(n1c)::value;
      static_for<1, R2>([&](auto k2c) {
        constexpr int k2 = decltype(k2c)::value;
        a[n1][k2] = twc<R, n1 * k2, SIGN>(a[n1][k2]);
      });
    });
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      cpx b[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) b[n1] = a[n1][k2];
      dft<R1, SIGN, CONJ_HI>(b);  // outputs k2 + R2 k1, k1 >= R1/2 <=> index >= R/2
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) x[k2 + R2 * k1] = b[k1];
    }
  }
}

// Padded LDS index: one pad slot per 32 elements breaks the power-of-two
// strides of the Stockham write pattern across LDS banks.
__host__ __device__ __forceinline__ int lds_pad(int n) { return n + (n >> 5); }
__host__ __device__ constexpr int lds_padded_size(int n) { return n + (n >> 5) + 1; }

}  // namespace spk
