// Batched LDS-resident mixed-radix FFT building blocks for gfx950.
//
// A transform of length L is a short sequence of Stockham stages with large
// radices (16, 24, 32, ...), so L = 384 takes two stages.  In a stage every
// task (vector v, butterfly j) holds R points x[j + t L/R] in registers,
// multiplies them by W_{Ns R}^{t k} (k = j mod Ns), runs an R-point DFT in
// registers (itself a compile-time Cooley-Tukey split with twiddles folded
// into instruction literals) and stores the results at their auto-sorted
// positions (j - k) R + k + t Ns: the output is in natural order.
//
// Tasks are mapped vector-fastest (vec = idx mod V), vectors sit in LDS at an
// odd stride VS: lanes of a wave address different vectors at fixed position
// offsets, which keeps the stages free of LDS bank conflicts when V >= 16.  Stages expose
// gather / compute / scatter separately so kernels can read the first stage
// straight from global memory, keep the last stage in registers (spectrum
// multiply, inverse first stage) and write results straight back.
//
// The LDS twiddle table ``tw`` holds exp(-2 pi i e / L), e < L.  Inverse
// transforms conjugate every twiddle and are unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace jubik {

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // a * conj(b)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
  return {a.re * s, a.im * s};
}
// Multiply by -i (forward) or +i (inverse).
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) {
  return INV ? cplx<T>{-a.im, a.re} : cplx<T>{a.im, -a.re};
}

// ------------------------------------------------- compile-time unit roots
// exp(-2 pi i e / N) from Taylor polynomials on [0, pi/4] after an exact
// octant reduction of e / N; evaluated by the compiler, so every constant
// twiddle of a register DFT becomes an instruction literal.
struct dpair {
  double re, im;
};
constexpr double kQuarterPi = 0.78539816339744830961566084581987572;
constexpr double poly_sin(double x) {
  double x2 = x * x, t = x, s = x;
  for (int i = 1; i < 13; ++i) {
    t *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
    s += t;
  }
  return s;
}
constexpr double poly_cos(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int i = 1; i < 13; ++i) {
    t *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
    s += t;
  }
  return s;
}
constexpr dpair unit_root(long e, long n) {
  const long r = ((e % n) + n) % n;
  const long o = (8 * r) / n;
  const long rr = 8 * r - o * n;
  const double phi = kQuarterPi * (double)rr / (double)n;
  const double psi = kQuarterPi * (double)(n - rr) / (double)n;
  double c = 0, s = 0;  // cos and sin of 2 pi r / n
  switch (o) {
    case 0: c = poly_cos(phi); s = poly_sin(phi); break;
    case 1: c = poly_sin(psi); s = poly_cos(psi); break;
    case 2: c = -poly_sin(phi); s = poly_cos(phi); break;
    case 3: c = -poly_cos(psi); s = poly_sin(psi); break;
    case 4: c = -poly_cos(phi); s = -poly_sin(phi); break;
    case 5: c = -poly_sin(psi); s = -poly_cos(psi); break;
    case 6: c = poly_sin(phi); s = -poly_cos(phi); break;
    default: c = poly_cos(psi); s = -poly_sin(psi); break;
  }
  return {c, -s};
}
template <int N>
struct RootTable {
  double re[N];
  double im[N];
  constexpr RootTable() : re{}, im{} {
    for (int e = 0; e < N; ++e) {
      const dpair w = unit_root(e, N);
      re[e] = w.re;
      im[e] = w.im;
    }
  }
};
template <int N>
inline constexpr RootTable<N> kRoots{};

// x * W_N^e (conjugated for the inverse); e is a constant after unrolling.
template <typename T, int N, bool INV>
__device__ __forceinline__ cplx<T> twiddle_c(cplx<T> x, int e) {
  if (e == 0) return x;
  if (4 * e == N) return mul_mi<T, INV>(x);
  if (2 * e == N) return {-x.re, -x.im};
  if (4 * e == 3 * N) return mul_mi<T, !INV>(x);
  const T wr = T(kRoots<N>.re[e]);
  const T wi = INV ? T(-kRoots<N>.im[e]) : T(kRoots<N>.im[e]);
  return {x.re * wr - x.im * wi, x.re * wi + x.im * wr};
}

// ---------------------------------------------------------------- small DFTs
template <typename T, int R, bool INV>
struct Dft;

template <typename T, bool INV>
struct Dft<T, 2, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T, bool INV>
struct Dft<T, 4, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    cplx<T> c = cadd(v[1], v[3]), d = mul_mi<T, INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

template <typename T, bool INV>
struct Dft<T, 8, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    const T h = T(0.70710678118654752440084436210485);
    cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, INV>::run(e);
    Dft<T, 4, INV>::run(o);
    cplx<T> o1 = INV ? cplx<T>{h * (o[1].re - o[1].im), h * (o[1].re + o[1].im)}
                     : cplx<T>{h * (o[1].re + o[1].im), h * (o[1].im - o[1].re)};
    cplx<T> o2 = mul_mi<T, INV>(o[2]);
    cplx<T> o3 = INV ? cplx<T>{-h * (o[3].re + o[3].im), h * (o[3].re - o[3].im)}
                     : cplx<T>{h * (o[3].im - o[3].re), -h * (o[3].re + o[3].im)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <typename T, bool INV>
struct Dft<T, 3, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    const T s3 = T(0.86602540378443864676372317075294);
    cplx<T> t1 = cadd(v[1], v[2]);
    cplx<T> t2 = cscale(csub(v[1], v[2]), s3);
    cplx<T> m = {v[0].re - T(0.5) * t1.re, v[0].im - T(0.5) * t1.im};
    cplx<T> r = mul_mi<T, INV>(t2);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <typename T, bool INV>
struct Dft<T, 5, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    const T c1 = T(0.30901699437494742410229341718282);
    const T c2 = T(-0.80901699437494742410229341718282);
    const T s1 = T(0.95105651629515357211643933337938);
    const T s2 = T(0.58778525229247312916870595463907);
    cplx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    cplx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    cplx<T> x0 = v[0];
    v[0] = cadd(x0, cadd(a1, a2));
    cplx<T> m1 = {x0.re + c1 * a1.re + c2 * a2.re, x0.im + c1 * a1.im + c2 * a2.im};
    cplx<T> m2 = {x0.re + c2 * a1.re + c1 * a2.re, x0.im + c2 * a1.im + c1 * a2.im};
    cplx<T> n1 = mul_mi<T, INV>(cplx<T>{s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im});
    cplx<T> n2 = mul_mi<T, INV>(cplx<T>{s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im});
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
};

template <typename T, bool INV>
struct Dft<T, 7, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    const T c1 = T(0.62348980185873353052500488400424);
    const T c2 = T(-0.22252093395631440428890256449679);
    const T c3 = T(-0.90096886790241912623610231950745);
    const T s1 = T(0.78183148246802980870844452667406);
    const T s2 = T(0.97492791218182360701813168299393);
    const T s3 = T(0.43388373911755812047576833284836);
    cplx<T> a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    cplx<T> a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    cplx<T> a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    cplx<T> x0 = v[0];
    v[0] = cadd(x0, cadd(a1, cadd(a2, a3)));
    // y_q = x0 + sum_t cos(2 pi t q / 7) a_t  -/+ i sum_t sin(2 pi t q / 7) b_t
    cplx<T> m1 = {x0.re + c1 * a1.re + c2 * a2.re + c3 * a3.re,
                  x0.im + c1 * a1.im + c2 * a2.im + c3 * a3.im};
    cplx<T> m2 = {x0.re + c2 * a1.re + c3 * a2.re + c1 * a3.re,
                  x0.im + c2 * a1.im + c3 * a2.im + c1 * a3.im};
    cplx<T> m3 = {x0.re + c3 * a1.re + c1 * a2.re + c2 * a3.re,
                  x0.im + c3 * a1.im + c1 * a2.im + c2 * a3.im};
    // sin(2 pi t q / 7) for (q,t): q=1: s1 s2 s3; q=2: s2 -s3 -s1; q=3: s3 -s1 s2
    cplx<T> n1 = mul_mi<T, INV>(cplx<T>{s1 * b1.re + s2 * b2.re + s3 * b3.re,
                                        s1 * b1.im + s2 * b2.im + s3 * b3.im});
    cplx<T> n2 = mul_mi<T, INV>(cplx<T>{s2 * b1.re - s3 * b2.re - s1 * b3.re,
                                        s2 * b1.im - s3 * b2.im - s1 * b3.im});
    cplx<T> n3 = mul_mi<T, INV>(cplx<T>{s3 * b1.re - s1 * b2.re + s2 * b3.re,
                                        s3 * b1.im - s1 * b2.im + s2 * b3.im});
    v[1] = cadd(m1, n1);
    v[6] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[5] = csub(m2, n2);
    v[3] = cadd(m3, n3);
    v[4] = csub(m3, n3);
  }
};

// Register DFT of any 7-smooth size: N = A * B with the inner B-point DFTs
// on the A strided subsequences, constant twiddles W_N^{i k}, and the outer
// A-point DFTs (output index k + B q).
__host__ __device__ constexpr bool dft_is_base(int N) {
  return N == 2 || N == 3 || N == 4 || N == 5 || N == 7 || N == 8;
}
__host__ __device__ constexpr int dft_split(int N) {
  return N % 4 == 0 ? 4 : N % 2 == 0 ? 2 : N % 3 == 0 ? 3 : N % 5 == 0 ? 5 : 7;
}

template <typename T, int N, bool INV, bool BASE = dft_is_base(N)>
struct RegDft;

template <typename T, int N, bool INV>
struct RegDft<T, N, INV, true> {
  static __device__ __forceinline__ void run(cplx<T>* v) { Dft<T, N, INV>::run(v); }
};

template <typename T, int N, bool INV>
struct RegDft<T, N, INV, false> {
  static constexpr int A = dft_split(N);
  static constexpr int B = N / A;
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> u[A][B];
#pragma unroll
    for (int i = 0; i < A; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) u[i][j] = v[i + A * j];
#pragma unroll
    for (int i = 0; i < A; ++i) RegDft<T, B, INV>::run(u[i]);
#pragma unroll
    for (int i = 1; i < A; ++i)
#pragma unroll
      for (int k = 1; k < B; ++k) u[i][k] = twiddle_c<T, N, INV>(u[i][k], i * k);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      cplx<T> w[A];
#pragma unroll
      for (int i = 0; i < A; ++i) w[i] = u[i][k];
      RegDft<T, A, INV>::run(w);
#pragma unroll
      for (int q = 0; q < A; ++q) v[k + B * q] = w[q];
    }
  }
};

// ---------------------------------------------------------------- plans
// Forward radix sequence per length; inverse transforms use it reversed, so
// the last forward stage and the first inverse stage share their register
// layout (tasks (vec, j) holding positions j + t L/R).
//
// Plan kind 0 takes the largest radices (fewest LDS passes; the column pass
// keeps its first and last stage in registers).  Plan kind 1 is for tiles of
// few vectors (V = 2 or 4 with 256 threads): smaller radices give every stage
// about one butterfly per thread, where large radices leave most waves idle
// while one wave runs the long butterflies.  Lengths without an entry use
// kind 0.
__host__ __device__ constexpr int plan_radix(int L, int s, int pl = 0) {
  if (pl == 1) {
    int w0 = 0, w1 = 0, w2 = 0;
    switch (L) {
      case 48: w0 = 4; w1 = 4; w2 = 3; break;
      case 64: w0 = 4; w1 = 4; w2 = 4; break;
      case 96: w0 = 4; w1 = 4; w2 = 6; break;
      case 128: w0 = 4; w1 = 4; w2 = 8; break;
      case 192: w0 = 8; w1 = 4; w2 = 6; break;
      case 224: w0 = 8; w1 = 4; w2 = 7; break;
      case 256: w0 = 4; w1 = 8; w2 = 8; break;
      case 320: w0 = 8; w1 = 8; w2 = 5; break;
      case 324: w0 = 6; w1 = 6; w2 = 9; break;
      case 384: w0 = 8; w1 = 8; w2 = 6; break;
      case 432: w0 = 8; w1 = 6; w2 = 9; break;
      case 567: w0 = 7; w1 = 9; w2 = 9; break;
      case 512: w0 = 8; w1 = 8; w2 = 8; break;
      case 1024: w0 = 8; w1 = 8; w2 = 16; break;
      case 1134: w0 = 14; w1 = 9; w2 = 9; break;
      default: break;
    }
    if (w0 > 0) return s == 0 ? w0 : s == 1 ? w1 : s == 2 ? w2 : 0;
  }
  int r0 = 0, r1 = 0, r2 = 0;
  switch (L) {
    case 2: r0 = 2; break;
    case 4: r0 = 4; break;
    case 8: r0 = 8; break;
    case 16: r0 = 16; break;
    case 24: r0 = 24; break;
    case 32: r0 = 32; break;
    case 48: r0 = 16; r1 = 3; break;
    case 64: r0 = 8; r1 = 8; break;
    case 96: r0 = 16; r1 = 6; break;
    case 128: r0 = 16; r1 = 8; break;
    case 192: r0 = 16; r1 = 12; break;
    case 224: r0 = 16; r1 = 14; break;
    case 256: r0 = 16; r1 = 16; break;
    case 320: r0 = 16; r1 = 20; break;
    case 324: r0 = 18; r1 = 18; break;
    case 384: r0 = 16; r1 = 24; break;
    case 432: r0 = 16; r1 = 27; break;
    case 512: r0 = 16; r1 = 32; break;
    case 567: r0 = 21; r1 = 27; break;
    case 768: r0 = 16; r1 = 16; r2 = 3; break;
    case 1024: r0 = 32; r1 = 32; break;
    case 1134: r0 = 27; r1 = 42; break;
    case 2048: r0 = 16; r1 = 16; r2 = 8; break;
    default: return -1;
  }
  return s == 0 ? r0 : s == 1 ? r1 : s == 2 ? r2 : 0;
}
__host__ __device__ constexpr int plan_stages(int L, int pl = 0) {
  int n = 0;
  while (n < 3 && plan_radix(L, n, pl) > 0) ++n;
  return n;
}
__host__ __device__ constexpr bool fft_length_ok(int L) { return plan_radix(L, 0) > 0; }
// Radix of stage s in execution order (reversed for the inverse).
__host__ __device__ constexpr int stage_radix(int L, bool inv, int s, int pl = 0) {
  const int n = plan_stages(L, pl);
  if (s >= n) return 0;
  return plan_radix(L, inv ? n - 1 - s : s, pl);
}
__host__ __device__ constexpr int stage_ns(int L, bool inv, int s, int pl = 0) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= stage_radix(L, inv, i, pl);
  return ns;
}
// Every plan multiplies out to its length.
__host__ __device__ constexpr bool plan_ok(int L, int pl) {
  int p = 1;
  for (int s = 0; s < plan_stages(L, pl); ++s) p *= plan_radix(L, s, pl);
  return p == L;
}
// Odd vector stride (in complex elements) for V vectors of length L.
__host__ __device__ constexpr int odd_stride(int L) { return L | 1; }
// LDS vector stride for a tile of V vectors with the vector-fastest task
// mapping: a 16-lane group of 8-byte accesses then holds min(V, 16) vectors
// times 16 / V consecutive positions, which land in distinct bank slots
// when the stride is congruent to 16 / V modulo 16 (odd for V >= 16).
__host__ __device__ constexpr int lds_stride(int L, int V) {
  const int r = V >= 16 ? 1 : 16 / V;
  return L + ((r - L % 16) % 16 + 16) % 16;
}

// ---------------------------------------------------------------- stages
template <typename T, int L, bool INV, int S, int V, int VS, int NT, int PL = 0>
struct Stage {
  static constexpr int R = stage_radix(L, INV, S, PL);
  static constexpr int Ns = stage_ns(L, INV, S, PL);
  static constexpr int NB = L / R;
  static constexpr int TOT = V * NB;
  static constexpr int BPT = (TOT + NT - 1) / NT;
  static constexpr int TWS = L / (Ns * R);
  static_assert((V & (V - 1)) == 0, "V must be a power of two");
  using Regs = cplx<T>[BPT][R];

  static __device__ __forceinline__ bool task(int tid, int b, int& vec, int& j) {
    const int idx = tid + b * NT;
    vec = idx & (V - 1);
    j = idx / V;
    return (TOT % NT == 0) || idx < TOT;
  }
  static __device__ __forceinline__ void gather_lds(const cplx<T>* buf, Regs& v, int tid) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      int vec, j;
      if (task(tid, b, vec, j)) {
        const cplx<T>* src = buf + vec * VS;
#pragma unroll
        for (int t = 0; t < R; ++t) v[b][t] = src[j + t * NB];
      }
    }
  }
  // v[b][t] = src(vec, j + t * NB)
  template <class Src>
  static __device__ __forceinline__ void gather(Regs& v, int tid, Src src) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      int vec, j;
      if (task(tid, b, vec, j)) {
#pragma unroll
        for (int t = 0; t < R; ++t) v[b][t] = src(vec, j + t * NB);
      }
    }
  }
  static __device__ __forceinline__ void compute(Regs& v, const cplx<T>* tw, int tid) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      int vec, j;
      if (task(tid, b, vec, j)) {
        if constexpr (Ns > 1) {
          const int k = j % Ns;
#pragma unroll
          for (int t = 1; t < R; ++t) {
            const cplx<T> w = tw[t * k * TWS];
            v[b][t] = INV ? cmulc(v[b][t], w) : cmul(v[b][t], w);
          }
        }
        RegDft<T, R, INV>::run(v[b]);
      }
    }
  }
  static __device__ __forceinline__ void scatter_lds(cplx<T>* buf, const Regs& v, int tid) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      int vec, j;
      if (task(tid, b, vec, j)) {
        const int k = j % Ns;
        cplx<T>* dst = buf + vec * VS;
        const int p0 = (j - k) * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) dst[p0 + t * Ns] = v[b][t];
      }
    }
  }
  // Last stage only (Ns * R == L): dst(vec, j + t * NB, value).
  template <class Dst>
  static __device__ __forceinline__ void emit(const Regs& v, int tid, Dst dst) {
    static_assert(Ns * R == L, "emit is for the last stage");
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      int vec, j;
      if (task(tid, b, vec, j)) {
#pragma unroll
        for (int t = 0; t < R; ++t) dst(vec, j + t * NB, v[b][t]);
      }
    }
  }
  // Apply f(vec, pos, value&) to every value held (pos = j + t NB for the
  // last stage, i.e. natural output positions).
  template <class F>
  static __device__ __forceinline__ void for_each(Regs& v, int tid, F f) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      int vec, j;
      if (task(tid, b, vec, j)) {
#pragma unroll
        for (int t = 0; t < R; ++t) f(vec, j + t * NB, v[b][t]);
      }
    }
  }
};

// Stages S..end, LDS -> LDS (a barrier after every scatter).
template <typename T, int L, bool INV, int S, int V, int VS, int NT, int PL = 0>
__device__ __forceinline__ void fft_stages_lds(cplx<T>* buf, const cplx<T>* tw, int tid) {
  if constexpr (S < plan_stages(L, PL)) {
    using St = Stage<T, L, INV, S, V, VS, NT, PL>;
    typename St::Regs v;
    St::gather_lds(buf, v, tid);
    St::compute(v, tw, tid);
    __syncthreads();
    St::scatter_lds(buf, v, tid);
    __syncthreads();
    fft_stages_lds<T, L, INV, S + 1, V, VS, NT, PL>(buf, tw, tid);
  }
}

// Whole transform of V vectors in LDS (stride VS) with plan kind PL.  Caller
// synchronises after filling ``buf`` and ``tw``; the routine ends with a
// barrier.
template <typename T, int L, int V, int VS, int NT, bool INV, int PL = 0>
__device__ __forceinline__ void fft_lds(cplx<T>* buf, const cplx<T>* tw, int tid) {
  static_assert(fft_length_ok(L), "FFT length has no plan");
  static_assert(plan_ok(L, PL), "plan radices do not multiply to the length");
  fft_stages_lds<T, L, INV, 0, V, VS, NT, PL>(buf, tw, tid);
}

// Plan kind for tiles of V vectors of length L on NT threads: kind 1 when
// the stage of the largest radix would have fewer butterflies than threads.
__host__ __device__ constexpr int plan_max_radix(int L, int pl) {
  int r = 0;
  for (int s = 0; s < plan_stages(L, pl); ++s) r = plan_radix(L, s, pl) > r ? plan_radix(L, s, pl) : r;
  return r;
}
template <int L, int V, int NT>
inline constexpr int kFewVectorPlan = (V * L / plan_max_radix(L, 0) < NT) ? 1 : 0;

template <typename T>
__device__ __forceinline__ void load_twiddles(cplx<T>* tw_lds, const cplx<T>* tw_glob, int L,
                                              int tid, int nt) {
  for (int e = tid; e < L; e += nt) tw_lds[e] = tw_glob[e];
}

}  // namespace jubik
