// Register/LDS Stockham FFT for one line of N = 2^n complex float32 points.
//
// Algorithm: Stockham autosort, decimation in time (Govindaraju, Lloyd,
// Dotsenko, Smith, Manferdelli, "High Performance Discrete Fourier Transforms
// on Graphics Processors", SC'08).  Stage s has radix R and stride Ns; the
// butterfly j reads x[j + r*N/R], twiddles by W_{Ns*R}^{r*(j mod Ns)}, and
// writes y[(j - j mod Ns)*R + j mod Ns + r*Ns].
//
// Data layout: a line is processed by T = N/E threads; thread t holds the E
// elements t + m*T (m = 0..E-1) in registers on entry and on exit, so the
// first stage's loads and the last stage's stores are unit-stride across a
// wavefront (coalesced when the line is contiguous, and C-wide segments for a
// tile of C columns).  Stages: one radix-R0 stage (R0 = N / E^S, no
// twiddles since Ns = 1) followed by S radix-E stages.  Between stages the
// data crosses threads through an LDS line (each stage with its own
// bank-conflict-free slot map, see Slot); the last stage is in place in
// registers.
#pragma once
#include "common.hpp"

namespace bpmfft {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// LDS slot of logical index p during the exchange after a stage of radix
// 2^RRL and stride 2^NSL.  LDS serves a wavefront's 64-bit accesses in groups
// of 16 lanes over 32 4-byte banks (the bank-conflict counters confirm this
// model on gfx950), so an access is conflict-free when the group's slots are
// distinct modulo 16.  A line whose threads fill a group needs its slots
// distinct in their low A = 4 bits; C lines sharing a group (column tiles:
// lane c of every C works on line c) each need A = 4 - log2(C) bits, with the
// line buffers offset by 16/C slots.  The map also makes the low A+1 bits
// distinct, for a 32-lane service group.
//
// Reads are contiguous (p = t + m*T) and are conflict-free for any map that
// XORs low bits with higher ones.  The scattered writes p = base + r*2^NSL
// put the radix digit r at bits [NSL, NSL+RRL) and push the lane bits
// j_NSL.. up by RRL, so the low A bits repeat across lanes.  The map XORs the
// displaced lane bits back into those positions.  The holes [NSL, min(NSL+RRL,
// A)) get bits from max(NSL+RRL, A) on, and bit A gets bit A+RRL when it does
// not already hold j_A.  Then the low A (and A+1) bits of the written slots
// are a bijection of the group's lane bits.  Every map is a permutation of
// [0, N): each output bit is XORed with a strictly higher input bit, so p is
// recovered from the top bit down.  The map is linear in the bits and its
// inputs lie above the radix digit, so Slot(base + r*2^NSL) = Slot(base) ^
// (r << NSL).
template <int A, int NSL, int RRL>
struct Slot {
  static constexpr int NH = NSL < A ? ((A - NSL) < RRL ? (A - NSL) : RRL) : 0;
  static constexpr int SRC = (NSL + RRL) > A ? (NSL + RRL) : A;
  static constexpr bool BITA = A >= 0 && A >= NSL;
  // the map leaves the radix digit's bits alone: r << NSL can be added
  static constexpr bool DIGIT_CLEAN = NH == 0 && !(BITA && A < NSL + RRL);
  __host__ __device__ static constexpr int map(int p) {
    int f = 0;
    if (NH > 0) f |= ((p >> SRC) & ((1 << NH) - 1)) << NSL;
    if (BITA) f |= ((p >> (A + RRL)) & 1) << A;
    return p ^ f;
  }
  // map permutes [0, N) (evaluated at compile time for every stage used)
  template <int N>
  static constexpr bool permutes() {
    bool seen[N] = {};
    for (int p = 0; p < N; ++p) {
      const int q = map(p);
      if (q < 0 || q >= N || seen[q]) return false;
      seen[q] = true;
    }
    return true;
  }
};

// Exchange layouts.  wr: offset in the line's LDS buffer of element
// base + r*2^NSL written after a stage of radix 2^RRL and stride 2^NSL (base
// has zero radix-digit bits); rd: offset of element t + m*T read back.
// Offsets are bytes when BYTES (a per-element XOR then needs no shift after
// it), else element indices (which let the compiler fold constant parts into
// the LDS instructions' offsets); each layout uses the form that compiles to
// fewer instructions for the kernels here.
//
// SwzLayout<A>: the Slot maps above (no padding); A = 4 for a line whose
// threads fill 16-lane groups (bytes), A = 4 - log2(C) when C lines share
// each group (element indices).
template <int A_ = 4>
struct SwzLayout {
  static constexpr int A = A_;
  static constexpr bool BYTES = A == 4;
  static constexpr int SH = BYTES ? 3 : 0;
  template <int N, int NSL, int RRL>
  static constexpr bool valid() { return Slot<A, NSL, RRL>::template permutes<N>(); }
  template <int NSL, int RRL>
  BPMH_DEV static int wr(int base, int r) {
    using S = Slot<A, NSL, RRL>;
    const int x = S::map(base) << SH;
    const int d = r << (NSL + SH);
    return S::DIGIT_CLEAN ? x + d : x ^ d;
  }
  // t < T and m*T share no bits, and the map is linear in the bits, so
  // Slot(t + m*T) = Slot(t) ^ Slot(m*T); a constant above T is added
  template <int NSL, int RRL, int T>
  BPMH_DEV static int rd(int t, int m) {
    using S = Slot<A, NSL, RRL>;
    const int y = S::map(t) << SH, km = S::map(m * T);
    return (km & (T - 1)) == 0 ? y + (km << SH) : y ^ (km << SH);
  }
};

// PadLayout: one pad slot after every 2^SH elements.
template <int SH>
struct PadLayout {
  static constexpr bool BYTES = false;
  template <int N, int NSL, int RRL>
  static constexpr bool valid() { return true; }
  template <int NSL, int RRL>
  BPMH_DEV static int wr(int base, int r) {
    const int p = base + (r << NSL);
    return p + (p >> SH);
  }
  template <int NSL, int RRL, int T>
  BPMH_DEV static int rd(int t, int m) {
    const int p = t + m * T;
    return p + (p >> SH);
  }
};

template <class L>
BPMH_DEV float2& lds_at(float2* lds, int off) {
  if constexpr (L::BYTES) return *reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + off);
  else return lds[off];
}

// s * 2^-1/2 with a two-term constant: the float32 rounding of 2^-1/2 alone
// biases every W8^odd product by -1.7e-8, which compounds over thousands of
// transforms; hi+lo keeps the product unbiased.
BPMH_DEV float mul_rsqrt2(float s) {
  const float hi = 0.70710677f;       // nearest float32 to 2^-1/2
  const float lo = 1.21016175e-08f;   // 2^-1/2 - hi
  return fmaf(s, lo, s * hi);
}

template <bool INV>
BPMH_DEV void dft(float2 (&a)[2]) {
  const float2 t = a[0];
  a[0] = c_add(t, a[1]);
  a[1] = c_sub(t, a[1]);
}

template <bool INV>
BPMH_DEV void dft4v(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 s02 = c_add(a0, a2), d02 = c_sub(a0, a2);
  const float2 s13 = c_add(a1, a3), d13 = c_rot90<INV>(c_sub(a1, a3));
  a0 = c_add(s02, s13);
  a2 = c_sub(s02, s13);
  a1 = c_add(d02, d13);
  a3 = c_sub(d02, d13);
}

template <bool INV>
BPMH_DEV void dft(float2 (&a)[4]) { dft4v<INV>(a[0], a[1], a[2], a[3]); }

template <bool INV>
BPMH_DEV void dft(float2 (&a)[8]) {
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4v<INV>(e0, e1, e2, e3);
  dft4v<INV>(o0, o1, o2, o3);
  if (!INV) {  // o_k *= exp(-2 pi i k/8)
    o1 = make_float2(mul_rsqrt2(o1.x + o1.y), mul_rsqrt2(o1.y - o1.x));
    o2 = c_rot90<false>(o2);
    o3 = make_float2(mul_rsqrt2(o3.y - o3.x), -mul_rsqrt2(o3.x + o3.y));
  } else {     // o_k *= exp(+2 pi i k/8)
    o1 = make_float2(mul_rsqrt2(o1.x - o1.y), mul_rsqrt2(o1.x + o1.y));
    o2 = c_rot90<true>(o2);
    o3 = make_float2(-mul_rsqrt2(o3.x + o3.y), mul_rsqrt2(o3.x - o3.y));
  }
  a[0] = c_add(e0, o0); a[4] = c_sub(e0, o0);
  a[1] = c_add(e1, o1); a[5] = c_sub(e1, o1);
  a[2] = c_add(e2, o2); a[6] = c_sub(e2, o2);
  a[3] = c_add(e3, o3); a[7] = c_sub(e3, o3);
}

// Orders the LDS exchange between the threads of one line.  WAVE: the line's
// threads are all in one wavefront, whose LDS operations complete in issue
// order, so a wavefront-scope fence (no s_barrier) suffices and the other
// wavefronts of the work-group run on undisturbed; otherwise a work-group
// barrier.
template <bool WAVE>
BPMH_DEV void line_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham stage.  w[r-1] = exp(-2 pi i r k/(NS R)), k = t mod NS, the
// thread's twiddles of this stage (only radix-E stages, B = 1, have NS > 1).
template <int N, int E, int R, int NS, bool INV, bool LAST, class L, bool WAVE>
BPMH_DEV void stage(float2 (&v)[E], float2* __restrict__ lds, int t,
                    const float2* __restrict__ w) {
  constexpr int T = N / E;
  constexpr int B = E / R;
  constexpr int NSL = ilog2c(NS), RRL = ilog2c(R);
  static_assert(NS == 1 || B == 1, "twiddled stages are full radix-E stages");
  static_assert(L::template valid<N, NSL, RRL>(), "exchange layout must permute the line");
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const int j = t + q * T;
    const int k = j & (NS - 1);
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[q + r * B];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = INV ? c_mulc(a[r], w[r - 1]) : c_mul(a[r], w[r - 1]);
    }
    dft<INV>(a);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * B] = a[r];
    } else {
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds_at<L>(lds, L::template wr<NSL, RRL>(base, r)) = a[r];
    }
  }
  if constexpr (!LAST) {
    line_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = lds_at<L>(lds, L::template rd<NSL, RRL, T>(t, m));
    line_sync<WAVE>();
  }
}

// Radix-E stages from stride NS on; w points at this stage's E-1 twiddles
// (none for NS == 1).
template <int N, int E, int NS, bool INV, class L, bool WAVE>
struct RadixEStages {
  BPMH_DEV static void run(float2 (&v)[E], float2* lds, int t, const float2* w) {
    constexpr bool LAST = (NS * E == N);
    stage<N, E, E, NS, INV, LAST, L, WAVE>(v, lds, t, w);
    if constexpr (!LAST)
      RadixEStages<N, E, NS * E, INV, L, WAVE>::run(v, lds, t, NS > 1 ? w + (E - 1) : w);
  }
};

template <int N, int E>
struct Shape {
  static constexpr int n = ilog2c(N);
  static constexpr int e = ilog2c(E);
  static constexpr int S = n / e;
  static constexpr int R0 = 1 << (n - S * e);
  static constexpr int T = N / E;
  static_assert((1 << n) == N, "N must be a power of two");
  static_assert((1 << e) == E && E >= 2 && E <= 8, "E must be 2, 4 or 8");
  static_assert(N >= E * E, "N must hold at least two radix-E stages");
};

// The twiddles one thread needs for a whole transform, loaded once into
// registers (they depend on t only) and shared by the forward and inverse
// transforms of a pass: (E-1) per radix-E stage with Ns > 1.
template <int N, int E>
struct Twiddles {
  using Sh = Shape<N, E>;
  static constexpr int NS0 = Sh::R0 > 1 ? Sh::R0 : E;   // first stage with Ns > 1
  static constexpr int NST = Sh::R0 > 1 ? Sh::S : Sh::S - 1;
  static constexpr int COUNT = NST * (E - 1);
  float2 w[COUNT > 0 ? COUNT : 1];

  // tab: the plan's stage-major table, for each stage with stride Ns the
  // entries W_{Ns*E}^{r*k} at [(r-1)*Ns + k] (r = 1..E-1, k < Ns), stages in
  // order.  One load instruction then reads consecutive k across the
  // wavefront (a few cache lines) instead of gathering W_N^{r*k*N/(Ns*E)}
  // from all over a natural-order table.
  BPMH_DEV void load(int t, const float2* __restrict__ tab) {
    int ns = NS0, off = 0;
#pragma unroll
    for (int s = 0; s < NST; ++s) {
      const int k = t & (ns - 1);
#pragma unroll
      for (int r = 1; r < E; ++r) w[s * (E - 1) + r - 1] = tab[off + (r - 1) * ns + k];
      off += (E - 1) * ns;
      ns *= E;
    }
  }
};

// Unnormalised DFT (INV=false: exp(-2 pi i jk/N); INV=true: exp(+2 pi i jk/N))
// of the line held as v[m] = x[t + m*T]; result in the same layout.
// lds: the line's exchange buffer in layout L (SwzLayout or PadLayout; the
// pointer carries the line's offset).  Every thread of the work-group must
// call it (it synchronises the group); WAVE = true when the line's T threads
// form one wavefront (or part of one) and no other thread touches its LDS
// line.
template <int N, int E, bool INV, class L, bool WAVE = false>
BPMH_DEV void line_fft(float2 (&v)[E], float2* __restrict__ lds, int t,
                       const Twiddles<N, E>& tw) {
  using Sh = Shape<N, E>;
  if constexpr (Sh::R0 > 1) {
    stage<N, E, Sh::R0, 1, INV, false, L, WAVE>(v, lds, t, nullptr);
    RadixEStages<N, E, Sh::R0, INV, L, WAVE>::run(v, lds, t, tw.w);
  } else {
    RadixEStages<N, E, 1, INV, L, WAVE>::run(v, lds, t, tw.w);
  }
}

}  // namespace bpmfft
