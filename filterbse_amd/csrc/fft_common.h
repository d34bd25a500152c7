// fft_common.h -- complex helpers and fixed-radix butterflies shared by the
// per-axis FFT passes (fft_kernels.hip) and the plane/slab passes
// (fft_tile_kernels.hip).  Sign convention is FFTW's: forward
// exp(-2 pi i jk/n), backward exp(+2 pi i jk/n); sgn = -1 forward.
#pragma once
#include "fbse_internal.h"

namespace fbse {

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C> __device__ __forceinline__ C cconj(C a) {
  a.y = -a.y;
  return a;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}
template <typename C> __device__ __forceinline__ C csub(C a, C b) {
  a.x -= b.x;
  a.y -= b.y;
  return a;
}
// multiply by s*i (s = +-1)
template <typename C, typename T> __device__ __forceinline__ C cmul_i(C a, T s) {
  C r;
  r.x = -s * a.y;
  r.y = s * a.x;
  return r;
}

// contribution of one projector record to its point:
// Y_m (proj d_so[m] + nl_proj d_nl[m]) summed over m (phamiltonian.c:140-444)
template <typename T>
__device__ __forceinline__ void load_record(const T *v, T (&r)[8]) {
  // the record (8 values) is 16-byte aligned: read it as 16-byte vectors
  if constexpr (sizeof(T) == 4) {
    const float4 *v4 = reinterpret_cast<const float4 *>(v);
    const float4 p0 = v4[0], p1 = v4[1];
    r[0] = p0.x, r[1] = p0.y, r[2] = p0.z, r[3] = p0.w, r[4] = p1.x, r[5] = p1.y, r[6] = p1.z, r[7] = p1.w;
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) r[q] = v[q];
  }
}

// One projector record's term of the expansion, sum_m Y_m (w_so d[m] + w_nl d[3+m]),
// for a record r[8] = (w_so, w_nl, Re/Im Y_m) and the 6 coefficients dd
// (16-byte aligned).
template <typename T>
__device__ __forceinline__ cplx_t<T> record_term_r(const T (&r)[8], const cplx_t<T> *dd) {
  cplx_t<T> d[6];
  if constexpr (sizeof(T) == 4) {
    const float4 *d4 = reinterpret_cast<const float4 *>(dd);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float4 t = d4[q];
      d[2 * q] = cplx_t<T>{t.x, t.y};
      d[2 * q + 1] = cplx_t<T>{t.z, t.w};
    }
  } else {
#pragma unroll
    for (int q = 0; q < 6; ++q) d[q] = dd[q];
  }
  cplx_t<T> acc{(T)0, (T)0};
  const T wso = r[0], wnl = r[1];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const cplx_t<T> a = d[m], b = d[3 + m];
    const T sx = wso * a.x + wnl * b.x, sy = wso * a.y + wnl * b.y;
    const T yx = r[2 + 2 * m], yy = r[3 + 2 * m];
    acc.x += yx * sx - yy * sy;
    acc.y += yx * sy + yy * sx;
  }
  return acc;
}

template <typename T>
__device__ __forceinline__ cplx_t<T> record_term(const T *v, const cplx_t<T> *dd) {
  T r[8];
  load_record<T>(v, r);
  return record_term_r<T>(r, dd);
}

// Butterfly of radix R on values x[0..R-1] (already twiddled); sgn = -1 forward.
// Output q is left in x[perm(q)] (the identity for the prime-power radices).
template <typename T, int R> struct Bfly;
struct NaturalOrder {
  __host__ __device__ static constexpr int perm(int q) { return q; }
};

template <typename T> struct Bfly<T, 2> : NaturalOrder {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C *x, T) {
    C a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};
template <typename T> struct Bfly<T, 3> : NaturalOrder {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C *x, T sgn) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    C t1 = cadd(x[1], x[2]);
    C t2;
    t2.x = x[0].x - (T)0.5 * t1.x;
    t2.y = x[0].y - (T)0.5 * t1.y;
    C d = csub(x[1], x[2]);
    C t3 = cmul_i<C, T>(d, sgn * h);
    x[0] = cadd(x[0], t1);
    x[1] = cadd(t2, t3);
    x[2] = csub(t2, t3);
  }
};
template <typename T> struct Bfly<T, 4> : NaturalOrder {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C *x, T sgn) {
    C a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    C c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
    C di = cmul_i<C, T>(d, sgn);
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = cadd(b, di);
    x[3] = csub(b, di);
  }
};
template <typename T> struct Bfly<T, 5> : NaturalOrder {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C *x, T sgn) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    C t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
    C t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
    C a1, a2, u1, u2;
    a1.x = x[0].x + c1 * t1.x + c2 * t2.x;
    a1.y = x[0].y + c1 * t1.y + c2 * t2.y;
    a2.x = x[0].x + c2 * t1.x + c1 * t2.x;
    a2.y = x[0].y + c2 * t1.y + c1 * t2.y;
    u1.x = s1 * t3.x + s2 * t4.x;
    u1.y = s1 * t3.y + s2 * t4.y;
    u2.x = s2 * t3.x - s1 * t4.x;
    u2.y = s2 * t3.y - s1 * t4.y;
    C b1 = cmul_i<C, T>(u1, sgn), b2 = cmul_i<C, T>(u2, sgn);
    x[0] = cadd(x[0], cadd(t1, t2));
    x[1] = cadd(a1, b1);
    x[4] = csub(a1, b1);
    x[2] = cadd(a2, b2);
    x[3] = csub(a2, b2);
  }
};
template <typename T> struct Bfly<T, 8> : NaturalOrder {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C *x, T sgn) {
    C e[4] = {x[0], x[2], x[4], x[6]};
    C o[4] = {x[1], x[3], x[5], x[7]};
    Bfly<T, 4>::run(e, sgn);
    Bfly<T, 4>::run(o, sgn);
    const T r = (T)0.70710678118654752440;
    // w^1 = r(1 + sgn i), w^2 = sgn i, w^3 = r(-1 + sgn i)
    C w1o, w2o, w3o;
    w1o.x = r * (o[1].x - sgn * o[1].y);
    w1o.y = r * (o[1].y + sgn * o[1].x);
    w2o = cmul_i<C, T>(o[2], sgn);
    w3o.x = r * (-o[3].x - sgn * o[3].y);
    w3o.y = r * (-o[3].y + sgn * o[3].x);
    x[0] = cadd(e[0], o[0]);
    x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], w1o);
    x[5] = csub(e[1], w1o);
    x[2] = cadd(e[2], w2o);
    x[6] = csub(e[2], w2o);
    x[3] = cadd(e[3], w3o);
    x[7] = csub(e[3], w3o);
  }
};

// Composite butterflies R = P * Q (one register pass instead of two LDS
// passes): x[Q n1 + n2] -> DFT_P over n1 -> twiddle w_R^(n2 k1) -> DFT_Q over
// n2 -> X[k1 + P k2].  RootTab<N> holds cos and sin of 2 pi j / N, j < N.
template <int N> struct RootTab;
template <> struct RootTab<6> {
  static constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
  static constexpr double s[6] = {0.0, 0.866025403784438646764, 0.866025403784438646764, 0.0, -0.866025403784438646764, -0.866025403784438646764};
};
template <> struct RootTab<9> {
  static constexpr double c[9] = {1.0, 0.766044443118978035202, 0.173648177666930348852, -0.5, -0.939692620785908384054, -0.939692620785908384054, -0.5, 0.173648177666930348852, 0.766044443118978035202};
  static constexpr double s[9] = {0.0, 0.642787609686539326323, 0.984807753012208059367, 0.866025403784438646764, 0.342020143325668733044, -0.342020143325668733044, -0.866025403784438646764, -0.984807753012208059367, -0.642787609686539326323};
};
template <> struct RootTab<10> {
  static constexpr double c[10] = {1.0, 0.809016994374947424102, 0.309016994374947424102, -0.309016994374947424102, -0.809016994374947424102, -1.0, -0.809016994374947424102, -0.309016994374947424102, 0.309016994374947424102, 0.809016994374947424102};
  static constexpr double s[10] = {0.0, 0.587785252292473129169, 0.951056516295153572116, 0.951056516295153572116, 0.587785252292473129169, 0.0, -0.587785252292473129169, -0.951056516295153572116, -0.951056516295153572116, -0.587785252292473129169};
};
template <> struct RootTab<12> {
  static constexpr double c[12] = {1.0, 0.866025403784438646764, 0.5, 0.0, -0.5, -0.866025403784438646764, -1.0, -0.866025403784438646764, -0.5, 0.0, 0.5, 0.866025403784438646764};
  static constexpr double s[12] = {0.0, 0.5, 0.866025403784438646764, 1.0, 0.866025403784438646764, 0.5, 0.0, -0.5, -0.866025403784438646764, -1.0, -0.866025403784438646764, -0.5};
};
template <> struct RootTab<15> {
  static constexpr double c[15] = {1.0, 0.913545457642600895502, 0.669130606358858213826, 0.309016994374947424102, -0.1045284632676534714, -0.5, -0.809016994374947424102, -0.978147600733805637929, -0.978147600733805637929, -0.809016994374947424102, -0.5, -0.1045284632676534714, 0.309016994374947424102, 0.669130606358858213826, 0.913545457642600895502};
  static constexpr double s[15] = {0.0, 0.406736643075800207754, 0.743144825477394235015, 0.951056516295153572116, 0.994521895368273336923, 0.866025403784438646764, 0.587785252292473129169, 0.207911690817759337102, -0.207911690817759337102, -0.587785252292473129169, -0.866025403784438646764, -0.994521895368273336923, -0.951056516295153572116, -0.743144825477394235015, -0.406736643075800207754};
};
template <> struct RootTab<16> {
  static constexpr double c[16] = {1.0, 0.923879532511286756128, 0.707106781186547524401, 0.382683432365089771728, 0.0, -0.382683432365089771728, -0.707106781186547524401, -0.923879532511286756128, -1.0, -0.923879532511286756128, -0.707106781186547524401, -0.382683432365089771728, 0.0, 0.382683432365089771728, 0.707106781186547524401, 0.923879532511286756128};
  static constexpr double s[16] = {0.0, 0.382683432365089771728, 0.707106781186547524401, 0.923879532511286756128, 1.0, 0.923879532511286756128, 0.707106781186547524401, 0.382683432365089771728, 0.0, -0.382683432365089771728, -0.707106781186547524401, -0.923879532511286756128, -1.0, -0.923879532511286756128, -0.707106781186547524401, -0.382683432365089771728};
};

template <typename T, int P, int Q> struct BflyPQ {
  using C = cplx_t<T>;
  // In place, output permuted: X[k1 + P k2] ends in x[perm(k1 + P k2)] =
  // x[Q k1 + k2] (compile-time indices, so the permutation costs nothing).
  __host__ __device__ static constexpr int perm(int q) { return Q * (q % P) + q / P; }
  __device__ __forceinline__ static void run(C *x, T sgn) {
    constexpr int N = P * Q;
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      C t[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) t[n1] = x[Q * n1 + n2];
      Bfly<T, P>::run(t, sgn);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        C v = t[k1];
        if (n2 > 0 && k1 > 0) {
          // after unrolling j is a constant: roots +-1 and +-i cost no multiply
          const int j = (n2 * k1) % N;
          if (RootTab<N>::c[j] == 0.0) {
            v = cmul_i<C, T>(v, sgn * (T)RootTab<N>::s[j]);
          } else if (RootTab<N>::s[j] == 0.0) {
            v.x = (T)RootTab<N>::c[j] * v.x;
            v.y = (T)RootTab<N>::c[j] * v.y;
          } else {
            C w;
            w.x = (T)RootTab<N>::c[j];
            w.y = sgn * (T)RootTab<N>::s[j];
            v = cmul(v, w);
          }
        }
        x[Q * k1 + n2] = v;  // y[n2][k1]
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) Bfly<T, Q>::run(x + Q * k1, sgn);
  }
};
template <typename T> struct Bfly<T, 6> : BflyPQ<T, 2, 3> {};
template <typename T> struct Bfly<T, 9> : BflyPQ<T, 3, 3> {};
template <typename T> struct Bfly<T, 10> : BflyPQ<T, 2, 5> {};
template <typename T> struct Bfly<T, 12> : BflyPQ<T, 4, 3> {};
template <typename T> struct Bfly<T, 15> : BflyPQ<T, 3, 5> {};
template <typename T> struct Bfly<T, 16> : BflyPQ<T, 4, 4> {};

// x[r] *= w^r (r = 1 .. R-1) with w = tw[kk*tstep] (conjugated when inv): w
// and w^4 are read from the length-n table, w^(4q + j) = w^(4q) w^j with
// w^(4q) a product of reads (at most three roundings per power).
template <typename T, int R>
__device__ __forceinline__ void twiddle_powers(cplx_t<T> *x, const cplx_t<T> *tw, int kk, int tstep, bool inv) {
  using C = cplx_t<T>;
  C w1 = tw[kk * tstep];
  if (inv) w1 = cconj(w1);
  const C w2 = cmul(w1, w1), w3 = cmul(w2, w1);
  x[1] = cmul(x[1], w1);
  if (R > 2) x[2] = cmul(x[2], w2);
  if (R > 3) x[3] = cmul(x[3], w3);
  if (R > 4) {
    C w4 = tw[4 * kk * tstep];
    if (inv) w4 = cconj(w4);
    C base = w4;
#pragma unroll
    for (int q = 1; 4 * q < R; ++q) {
      if (q > 1) base = cmul(base, w4);
      x[4 * q] = cmul(x[4 * q], base);
      if (4 * q + 1 < R) x[4 * q + 1] = cmul(x[4 * q + 1], cmul(base, w1));
      if (4 * q + 2 < R) x[4 * q + 2] = cmul(x[4 * q + 2], cmul(base, w2));
      if (4 * q + 3 < R) x[4 * q + 3] = cmul(x[4 * q + 3], cmul(base, w3));
    }
  }
}

}  // namespace fbse
