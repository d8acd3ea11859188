// Register FFT codelets with compile-time sizes and twiddles (gfx950).
//
// Dft<T, N, SIGN>::run(v) transforms v[0..N-1] in registers: X[k] = sum_j v[j] e^{SIGN 2 pi i j k / N}
// (SIGN = -1 forward, +1 inverse, unnormalised). Composite sizes use the two-step split
// N = N1 * N2: X[k1 + N1 k2] = sum_{j2} e^{SIGN 2 pi i j2 k1 / N} e^{SIGN 2 pi i j2 k2 / N2}
// sum_{j1} v[j2 + N2 j1] e^{SIGN 2 pi i j1 k1 / N1}, recursively, with every twiddle a compile-time
// constant (evaluated by constexpr functions on template integers, so nothing is computed at run
// time and no table is read).
//
// LineSplit<N> gives the split used when a line of length N is shared by several threads: step 1
// runs N2 threads with a Dft<N1> each, step 2 runs N1 threads with a Dft<N2> each.
#pragma once
#include "nlfh_common.h"

namespace nlfh {
namespace fft {

constexpr double kPi = 3.14159265358979323846264338327950288;

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
// cos and sin of 2 pi k / n, reduced exactly to |phi| <= pi/4 around the nearest multiple of pi/2
constexpr double cos2pi(long k, long n) {
  k = ((k % n) + n) % n;
  const long q = (4 * k + n / 2) / n;  // nearest quarter turn
  const double phi = kPi * (double)(4 * k - q * n) / (2.0 * (double)n);
  const double c = taylor_cos(phi), s = taylor_sin(phi);
  switch (q & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
  }
}
constexpr double sin2pi(long k, long n) {
  k = ((k % n) + n) % n;
  const long q = (4 * k + n / 2) / n;
  const double phi = kPi * (double)(4 * k - q * n) / (2.0 * (double)n);
  const double c = taylor_cos(phi), s = taylor_sin(phi);
  switch (q & 3) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
  }
}

template <int N>
constexpr int first_factor() {
  return (N % 4 == 0 && N != 8) ? 4 : (N % 2 == 0 ? 2 : (N % 3 == 0 ? 3 : (N % 5 == 0 ? 5 : N)));
}

// compile-time loop
template <int I, int E>
struct For {
  template <class F>
  __host__ __device__ __forceinline__ static void run(F&& f) {
    f(std::integral_constant<int, I>{});
    For<I + 1, E>::run(f);
  }
};
template <int E>
struct For<E, E> {
  template <class F>
  __host__ __device__ __forceinline__ static void run(F&&) {}
};

// v * e^{SIGN 2 pi i K / N} with K, N compile-time
template <typename T, long K, long N, int SIGN>
__host__ __device__ __forceinline__ cplx<T> twiddle(cplx<T> v) {
  constexpr double c = cos2pi(K, N);
  constexpr double s = SIGN * sin2pi(K, N);
  if constexpr (K % N == 0) {
    return v;
  } else {
    return {v.re * (T)c - v.im * (T)s, v.re * (T)s + v.im * (T)c};
  }
}

template <typename T, int N, int SIGN>
struct Dft;

template <typename T, int SIGN>
struct Dft<T, 1, SIGN> {
  __host__ __device__ __forceinline__ static void run(cplx<T>*) {}
};

template <typename T, int SIGN>
struct Dft<T, 2, SIGN> {
  __host__ __device__ __forceinline__ static void run(cplx<T>* v) {
    const cplx<T> a = v[0], b = v[1];
    v[0] = {a.re + b.re, a.im + b.im};
    v[1] = {a.re - b.re, a.im - b.im};
  }
};

template <typename T, int SIGN>
struct Dft<T, 3, SIGN> {
  __host__ __device__ __forceinline__ static void run(cplx<T>* v) {
    const T s1 = (T)(SIGN * 0.86602540378443864676);
    const cplx<T> a0 = v[0], a1 = v[1], a2 = v[2];
    const T sr = a1.re + a2.re, si = a1.im + a2.im;
    const T dr = a1.re - a2.re, di = a1.im - a2.im;
    const T mr = a0.re - (T)0.5 * sr, mi = a0.im - (T)0.5 * si;
    v[0] = {a0.re + sr, a0.im + si};
    v[1] = {mr - s1 * di, mi + s1 * dr};
    v[2] = {mr + s1 * di, mi - s1 * dr};
  }
};

template <typename T, int SIGN>
struct Dft<T, 4, SIGN> {
  __host__ __device__ __forceinline__ static void run(cplx<T>* v) {
    const cplx<T> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
    const cplx<T> t0 = {a0.re + a2.re, a0.im + a2.im}, t1 = {a0.re - a2.re, a0.im - a2.im};
    const cplx<T> t2 = {a1.re + a3.re, a1.im + a3.im}, t3 = {a1.re - a3.re, a1.im - a3.im};
    // SIGN * i * t3
    const cplx<T> r3 = {(T)(-SIGN) * t3.im, (T)SIGN * t3.re};
    v[0] = {t0.re + t2.re, t0.im + t2.im};
    v[2] = {t0.re - t2.re, t0.im - t2.im};
    v[1] = {t1.re + r3.re, t1.im + r3.im};
    v[3] = {t1.re - r3.re, t1.im - r3.im};
  }
};

template <typename T, int SIGN>
struct Dft<T, 5, SIGN> {
  __host__ __device__ __forceinline__ static void run(cplx<T>* v) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)(SIGN * 0.95105651629515357212), s2 = (T)(SIGN * 0.58778525229247312917);
    const cplx<T> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3], a4 = v[4];
    const cplx<T> b1 = {a1.re + a4.re, a1.im + a4.im}, b4 = {a1.re - a4.re, a1.im - a4.im};
    const cplx<T> b2 = {a2.re + a3.re, a2.im + a3.im}, b3 = {a2.re - a3.re, a2.im - a3.im};
    const cplx<T> m1 = {a0.re + c1 * b1.re + c2 * b2.re, a0.im + c1 * b1.im + c2 * b2.im};
    const cplx<T> m2 = {a0.re + c2 * b1.re + c1 * b2.re, a0.im + c2 * b1.im + c1 * b2.im};
    const cplx<T> n1 = {s1 * b4.re + s2 * b3.re, s1 * b4.im + s2 * b3.im};
    const cplx<T> n2 = {s2 * b4.re - s1 * b3.re, s2 * b4.im - s1 * b3.im};
    v[0] = {a0.re + b1.re + b2.re, a0.im + b1.im + b2.im};
    v[1] = {m1.re - n1.im, m1.im + n1.re};
    v[4] = {m1.re + n1.im, m1.im - n1.re};
    v[2] = {m2.re - n2.im, m2.im + n2.re};
    v[3] = {m2.re + n2.im, m2.im - n2.re};
  }
};

// composite sizes: two-step split with compile-time twiddles
template <typename T, int N, int SIGN>
struct Dft {
  static constexpr int N1 = first_factor<N>() == N ? N : first_factor<N>();
  static constexpr int N2 = N / N1;
  static_assert(N1 > 1 && N2 > 1, "unsupported codelet size");
  __host__ __device__ __forceinline__ static void run(cplx<T>* v) {
    cplx<T> y[N];
    // step 1: N2 transforms of length N1 over v[j2 + N2 j1]; twiddle e^{SIGN 2 pi i j2 k1 / N}
    For<0, N2>::run([&](auto J2) {
      constexpr int j2 = decltype(J2)::value;
      cplx<T> s[N1];
      For<0, N1>::run([&](auto J1) { s[decltype(J1)::value] = v[j2 + N2 * decltype(J1)::value]; });
      Dft<T, N1, SIGN>::run(s);
      For<0, N1>::run([&](auto K1) {
        constexpr int k1 = decltype(K1)::value;
        y[j2 * N1 + k1] = twiddle<T, (long)j2 * k1, N, SIGN>(s[k1]);
      });
    });
    // step 2: N1 transforms of length N2 over j2 -> X[k1 + N1 k2]
    For<0, N1>::run([&](auto K1) {
      constexpr int k1 = decltype(K1)::value;
      cplx<T> s[N2];
      For<0, N2>::run([&](auto J2) { s[decltype(J2)::value] = y[decltype(J2)::value * N1 + k1]; });
      Dft<T, N2, SIGN>::run(s);
      For<0, N2>::run([&](auto K2) { v[k1 + N1 * decltype(K2)::value] = s[decltype(K2)::value]; });
    });
  }
};

// Split of a line of length N over threads (step 1: N2 threads x Dft<N1>; step 2: N1 threads x
// Dft<N2>); L = threads per line.
template <int N>
struct LineSplit;
template <>
struct LineSplit<16> { static constexpr int N1 = 4, N2 = 4; };
template <>
struct LineSplit<32> { static constexpr int N1 = 8, N2 = 4; };
template <>
struct LineSplit<48> { static constexpr int N1 = 8, N2 = 6; };
template <>
struct LineSplit<64> { static constexpr int N1 = 8, N2 = 8; };
template <>
struct LineSplit<128> { static constexpr int N1 = 16, N2 = 8; };

}  // namespace fft
}  // namespace nlfh
