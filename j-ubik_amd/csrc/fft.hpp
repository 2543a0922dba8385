// Batched LDS-resident mixed-radix FFT building blocks for gfx950.
//
// A work-group of NT threads transforms V complex vectors of length L that
// sit in LDS (vector v at buf + v * VS).  Each radix stage is a Stockham
// step: every thread loads R points (stride L/R) of a butterfly into
// registers, applies the twiddles W_{Ns R}^{t k}, runs the R-point DFT and,
// after a barrier, stores the results at their auto-sorted positions, so
// the output is in natural order and no bit reversal pass is needed.  L is
// a compile-time constant; its radix sequence is chosen greedily
// (8, then 4, 2, 3, 5, 7), so every 7-smooth length is supported.
//
// The twiddle table ``tw`` (LDS) holds exp(-2 pi i e / L), e = 0..L-1.
// Inverse transforms are unnormalised.
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

// ---------------------------------------------------------------- radix plan
__host__ __device__ constexpr int radix_at(int L, int i) {
  int r = 1;
  for (int s = 0; s <= i; ++s) {
    if (L % 8 == 0) r = 8;
    else if (L % 4 == 0) r = 4;
    else if (L % 2 == 0) r = 2;
    else if (L % 3 == 0) r = 3;
    else if (L % 5 == 0) r = 5;
    else if (L % 7 == 0) r = 7;
    else return (L == 1) ? 1 : -1;
    L /= r;
  }
  return r;
}
__host__ __device__ constexpr int num_stages(int L) {
  int s = 0;
  while (L > 1) {
    int r = radix_at(L, 0);
    if (r < 0) return -1;
    L /= r;
    ++s;
  }
  return s;
}
__host__ __device__ constexpr bool fft_length_ok(int L) { return L >= 2 && num_stages(L) > 0; }

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
    // o[q] *= W8^q (conjugated for the inverse)
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

// ---------------------------------------------------------------- one stage
template <typename T, int L, int R, int Ns, int V, int VS, int NT, bool INV>
__device__ __forceinline__ void fft_stage(cplx<T>* buf, const cplx<T>* tw, int tid) {
  constexpr int NB = L / R;
  constexpr int TOT = V * NB;
  constexpr int BPT = (TOT + NT - 1) / NT;
  constexpr int TWS = L / (Ns * R);
  cplx<T> v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int idx = tid + b * NT;
    if (TOT % NT == 0 || idx < TOT) {
      const int vec = idx / NB;
      const int j = idx - vec * NB;
      const cplx<T>* src = buf + vec * VS + j;
#pragma unroll
      for (int t = 0; t < R; ++t) v[b][t] = src[t * NB];
      if constexpr (Ns > 1) {
        const int k = j % Ns;
#pragma unroll
        for (int t = 1; t < R; ++t) {
          cplx<T> w = tw[t * k * TWS];
          v[b][t] = INV ? cmulc(v[b][t], w) : cmul(v[b][t], w);
        }
      }
      Dft<T, R, INV>::run(v[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int idx = tid + b * NT;
    if (TOT % NT == 0 || idx < TOT) {
      const int vec = idx / NB;
      const int j = idx - vec * NB;
      const int k = j % Ns;
      cplx<T>* dst = buf + vec * VS + (j - k) * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) dst[t * Ns] = v[b][t];
    }
  }
  __syncthreads();
}

template <typename T, int L, int Lrem, int Ns, int V, int VS, int NT, bool INV>
struct FftStages {
  static __device__ __forceinline__ void run(cplx<T>* buf, const cplx<T>* tw, int tid) {
    constexpr int R = radix_at(Lrem, 0);
    fft_stage<T, L, R, Ns, V, VS, NT, INV>(buf, tw, tid);
    FftStages<T, L, Lrem / R, Ns * R, V, VS, NT, INV>::run(buf, tw, tid);
  }
};
template <typename T, int L, int Ns, int V, int VS, int NT, bool INV>
struct FftStages<T, L, 1, Ns, V, VS, NT, INV> {
  static __device__ __forceinline__ void run(cplx<T>*, const cplx<T>*, int) {}
};

// Transform V vectors (stride VS) in LDS.  Caller must __syncthreads()
// after filling ``buf`` and ``tw``; the routine ends with a barrier.
template <typename T, int L, int V, int VS, int NT, bool INV>
__device__ __forceinline__ void fft_lds(cplx<T>* buf, const cplx<T>* tw, int tid) {
  static_assert(fft_length_ok(L), "FFT length must be 7-smooth");
  FftStages<T, L, L, 1, V, VS, NT, INV>::run(buf, tw, tid);
}

template <typename T>
__device__ __forceinline__ void load_twiddles(cplx<T>* tw_lds, const cplx<T>* tw_glob, int L,
                                              int tid, int nt) {
  for (int e = tid; e < L; e += nt) tw_lds[e] = tw_glob[e];
}

}  // namespace jubik
