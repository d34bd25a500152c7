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

// Butterfly of radix R on values x[0..R-1] (already twiddled); sgn = -1 forward.
template <typename T, int R> struct Bfly;

template <typename T> struct Bfly<T, 2> {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C *x, T) {
    C a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};
template <typename T> struct Bfly<T, 3> {
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
template <typename T> struct Bfly<T, 4> {
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
template <typename T> struct Bfly<T, 5> {
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
template <typename T> struct Bfly<T, 8> {
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

}  // namespace fbse
