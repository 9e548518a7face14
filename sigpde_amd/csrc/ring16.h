// Goursat PDE wavefront for lane groups of 16 (four pairs per wave) with a skewed LDS ring.
//
// Same scheme, cell arithmetic and virtual-column stream as ring_solver.h (stojl/SigPDE
// sigpde/cuda_device_functions.py:3-5, sigpde/cuda_pairwise_kernels.py:8-39): lane rho of a group
// owns refined row 16*b + rho of band b and handles virtual column v at step t = v + rho; the
// value above arrives from lane rho-1 by DPP (row_shr:1); bands of a pair and pairs of a stream
// follow each other as one stream of virtual columns, Np per band.
//
// What differs is the bookkeeping, laid out so that the hot loop has no address arithmetic:
//   * Skewed ring. Increment (row r, virtual column v) of the wave's 64 ring rows (4 pairs x 16)
//     lives in slot (v + (r & 15)) mod 32, element r of that slot (slots of kR16Stride elements).
//     At step t every lane reads slot t mod 32, element `lane`: one address per lane for the whole
//     solve, the slot an immediate offset. Chunk k (16 columns) is written in period k-1, after
//     the reads of that period were issued (its slots held chunk k-2, last read at step 16k-2).
//     Ring and boundary inputs are loaded one group (4 steps) before their use.
//   * Period loop unrolled by two (even / odd parity), so ring slots are compile-time constants.
//   * Band boundary line indexed by step: the group's last lane writes its value of step T at
//     index T mod Np; the group's first lane reads, at step T, index (T + 15) mod Np, written at
//     step T + 15 - Np (the previous band's value in the same column). Indices Np..Np+14 mirror
//     0..14 (written in the periods that start a band), so a period's reads never wrap.
//   * Three period flavours chosen once per period: Plain, Ones (the boundary values written are
//     ones: the group's last lane is in a pair's last band, whose next band is the next pair's
//     first), Special (a reset column, where k = 1, and/or the result cell).
//   * R refined rows per lane (R = min(2^order, 8)): lane rho owns rows R rho .. R rho + R - 1 of a
//     band of 16 R rows; at dyadic order >= log2 R these rows refine one original row, so at every
//     column all R cells share one increment: the coefficients and the DPP shift are paid once per
//     R cells. The lane's rows are updated top to bottom within a step (row r takes row r-1's new
//     value from the same lane); the ring holds one increment per lane and column, as for R = 1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ring_solver.h"

namespace sigpde {

constexpr int kR16Slots = 32;   // ring slots (skew 16 + one chunk of 16)

// Refined rows per lane at dyadic order o: 2^o, at most 8.
__host__ __device__ inline int r16_rows(int order) { return order >= 3 ? 8 : 1 << order; }
constexpr int kR16Stride = 67;  // elements per slot: 64 rows + 3 pad (conflict-free chunk writes)

// Element type of the band boundary line: float where every cell is rounded to float anyway
// (Ref32, Scaled32: exact), and for Mixed, whose band boundary rows are rounded to float (about
// 1e-7 relative on the result at 255 x 255 grids, far below the fp32 MFMA increment error); the
// float64 modes keep double.
template <Arith A>
struct R16Bnd {
  using T = typename std::conditional<A == Arith::Ref64 || A == Arith::Scaled64, double,
                                      float>::type;
};

// Arith::MixedF (the float32 wavefront, ring16f.h) lays its ring out by row instead: ring row r
// (the consumer lane) holds slots 0..31 at floats r * kR16FRowStride + slot, so that a lane's four
// columns of a consumer group are two 8-byte LDS reads; its lines carry a one-element shift (reads
// 16-byte aligned) and a block of kR16FZeros zero floats follows the four lines.
constexpr int kR16FRowStride = 34;  // 32 slots + 2 pad: 8-byte reads of 16 lanes hit 32 banks
constexpr int kR16FZeros = 16;

// Boundary line stride (elements): Np + 16 (mirror) plus a pad that puts the four lines of a wave
// in different LDS banks (the lane groups read and write them in the same instructions). MixedF:
// Np + 16 (mirror) + 1 (shift) rounded to a multiple of four, so every line starts 16-byte aligned.
template <Arith A>
__host__ __device__ inline int r16_line_stride(int Np) {
  if constexpr (A == Arith::MixedF) return Np + 20;
  return Np + 16 + (sizeof(typename R16Bnd<A>::T) == 4 ? 8 : 4);
}

// LDS of one wave whose boundary lines live in global memory (long rows of columns): the ring.
template <Arith A>
__host__ __device__ inline size_t r16_ring_bytes() {
  if constexpr (A == Arith::MixedF) return (size_t)kWave * kR16FRowStride * sizeof(float);
  return (size_t)kR16Slots * kR16Stride * sizeof(typename ArithTraits<A>::RT);
}

// Global memory for the four boundary lines of one wave (same layout as in LDS; MixedF: and the
// zero block).
template <Arith A>
__host__ __device__ inline size_t r16_line_bytes(int N) {
  using BT = typename R16Bnd<A>::T;
  return (size_t)4 * r16_line_stride<A>(Layout<16>::band_len(N)) * sizeof(BT) +
         (A == Arith::MixedF ? (size_t)kR16FZeros * sizeof(float) : 0);
}

// LDS for one wave: four boundary lines (MixedF: and the zero block), then the ring (the lines'
// total is a multiple of 64 bytes, so the ring stays aligned).
template <Arith A>
__host__ __device__ inline size_t r16_lds_bytes(int N) {
  return r16_line_bytes<A>(N) + r16_ring_bytes<A>();
}

// Chunk writes of a producer. The producing lane l = (c, h), c = l & 15, h = l >> 4, holds element
// e (0..3) of pair g: ring row 16g + 4h + e, chunk column c. Chunk k goes to slot
// (16 (k & 1) + c + 4h + e) & 31. Consecutive e are consecutive slots, so one base offset per
// parity serves all four elements (immediate offsets e * (stride + 1)); only odd chunks can wrap
// within a lane's four elements (when (16 + c + 4h) & 31 >= 29), corrected by one select each.
template <typename RT>
struct Ring16Writer {
  uint32_t base[2];  // byte offset of element 0 (pair 0) for even / odd chunks
  int s1;            // (16 + c + 4h) & 31: the odd chunks' first slot
  __device__ __forceinline__ void init() {
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, h = lane >> 4;
    const int s0 = c + 4 * h;  // <= 27: even chunks never wrap
    s1 = (16 + c + 4 * h) & 31;
    base[0] = (uint32_t)((s0 * kR16Stride + 4 * h) * sizeof(RT));
    base[1] = (uint32_t)((s1 * kR16Stride + 4 * h) * sizeof(RT));
  }
  // v[g][e]: the lane's four values of pair g
  template <int PAR>
  __device__ __forceinline__ void store(RT *ring, const RT (&v)[4][4]) const {
    char *rbase = reinterpret_cast<char *>(ring);
    constexpr uint32_t kWrap = (uint32_t)(kR16Slots * kR16Stride * sizeof(RT));
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint32_t a = base[PAR] + (uint32_t)(e * (kR16Stride + 1) * sizeof(RT));
      if (PAR == 1 && e > 0) a = s1 + e >= kR16Slots ? a - kWrap : a;
#pragma unroll
      for (int g = 0; g < 4; ++g) *reinterpret_cast<RT *>(rbase + a + g * 16 * sizeof(RT)) = v[g][e];
    }
  }
};

// Chunk writes into the row-major ring of Arith::MixedF: ring row r at floats r * kR16FRowStride,
// slot (16 (k & 1) + c + (r & 15)) & 31 as above. Element e of pair g sits at row 16g + 4h + e and
// slot s + e, i.e. e * (kR16FRowStride + 1) floats past element 0; the pairs are 16 rows apart.
struct Ring16WriterF {
  uint32_t base[2];  // byte offset of element 0 (pair 0) for even / odd chunks
  int s1;            // (16 + c + 4h) & 31: the odd chunks' first slot
  __device__ __forceinline__ void init() {
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, h = lane >> 4;
    const int s0 = c + 4 * h;  // <= 27: even chunks never wrap
    s1 = (16 + c + 4 * h) & 31;
    base[0] = (uint32_t)((4 * h * kR16FRowStride + s0) * sizeof(float));
    base[1] = (uint32_t)((4 * h * kR16FRowStride + s1) * sizeof(float));
  }
  template <int PAR>
  __device__ __forceinline__ void store(float *ring, const float (&v)[4][4]) const {
    char *rbase = reinterpret_cast<char *>(ring);
    constexpr uint32_t kWrap = (uint32_t)(kR16Slots * sizeof(float));
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint32_t a = base[PAR] + (uint32_t)(e * (kR16FRowStride + 1) * sizeof(float));
      if (PAR == 1 && e > 0) a = s1 + e >= kR16Slots ? a - kWrap : a;
#pragma unroll
      for (int g = 0; g < 4; ++g)
        *reinterpret_cast<float *>(rbase + a + g * 16 * kR16FRowStride * sizeof(float)) = v[g][e];
    }
  }
};

// Producer of increments read from global memory (any static kernel; torch computed them).
// Group g serves the pair whose increments (m x n, row stride ld) start at inc[g] (null: an idle
// slot, served zeros). Lane (c, h) loads element e of pair g = ring row 16g + 4h + e (the lane
// 4h + e of that group: refined rows R (16 band + 4h + e) ..) at chunk column c: four 64-byte
// row segments per load. Chunk k is loaded two periods before its use and scaled when stored.
// Plain (cached) loads: the other half of each 128-byte line is the next chunk's, read one period
// later from L2 (streaming loads that skip L2 refetched it from HBM: 25% slower).
template <typename RT>
struct Global16Producer {
  const RT *inc[4];
  RT scale[4];
  int64_t ld;
  bool scaled;
  RT buf[4][4];
  Cursor<16> cl;  // next chunk to load
  Ring16Writer<RT> W;

  __device__ __forceinline__ void load(const PairGeom<16> &g) {
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, h = lane >> 4;
    int jo = (cl.jv0 + c - 1) >> g.order;
    jo = jo < 0 ? 0 : (jo > g.n - 1 ? g.n - 1 : jo);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int ro = (cl.band * 16 * g.R + g.R * (4 * h + e)) >> g.order;
      ro = ro > g.m - 1 ? g.m - 1 : ro;
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const RT *base = inc[p];
        buf[p][e] = base ? base[(int64_t)ro * ld + jo] : (RT)0;
      }
    }
    cl.advance(g);
  }
  template <int PAR>
  __device__ __forceinline__ void store(RT *ring) {
    RT v[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int e = 0; e < 4; ++e) v[p][e] = scaled ? buf[p][e] * scale[p] : buf[p][e];
    W.template store<PAR>(ring, v);
  }
  __device__ __forceinline__ void prologue(RT *ring, const PairGeom<16> &g) {
    W.init();
    cl.start(g, 0);
    load(g);
    store<0>(ring);
    if (g.columns() > kChunk) load(g);
  }
  __device__ __forceinline__ void early(RT *, const PairGeom<16> &, int) {}
  // between consumer groups 2 and 3 of period per: write chunk per+1 (parity PAR), load per+2
  template <int PAR>
  __device__ __forceinline__ void period(RT *ring, const PairGeom<16> &g, int per) {
    const int nchunks = g.columns() / kChunk;
    if (per + 1 < nchunks) store<PAR>(ring);
    if (per + 2 < nchunks) load(g);
  }
  __device__ __forceinline__ double lane_scale() const { return 1.0; }
};

// Inputs of one consumer group (4 steps): ring values and boundary values.
template <Arith A>
struct R16In {
  typename ArithTraits<A>::RT c[4];
  typename R16Bnd<A>::T bv[4];
};

template <Arith A, int R>
struct R16State {
  using VT = typename ArithTraits<A>::VT;
  VT k[R];    // the lane's rows at its current column
  VT kPrev;   // the row above the lane's first row, at the previous column
  VT h, q;    // Mixed, MixedF: coefficients of the lane's current pair
  VT hN, qN;  // Mixed, MixedF: coefficients taken at the lane's next reset
};

template <bool R, bool O, bool C>
struct R16Flavour {
  static constexpr bool kReset = R;    // a reset column may fall in the period (runtime ones)
  static constexpr bool kOnes = O;     // the boundary values written are ones
  static constexpr bool kCapture = C;  // the result cell may fall in the period
};
using R16Plain = R16Flavour<false, false, false>;
using R16Ones = R16Flavour<false, true, false>;
using R16Special = R16Flavour<true, false, true>;

template <Arith A>
__device__ __forceinline__ void r16_load(const char *rb, int slot0,
                                         const typename R16Bnd<A>::T *bp, R16In<A> &in) {
  using RT = typename ArithTraits<A>::RT;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    in.c[q] = *reinterpret_cast<const RT *>(rb + (slot0 + q) * kR16Stride * sizeof(RT));
#pragma unroll
  for (int q = 0; q < 4; ++q) in.bv[q] = bp[q];
}

// Per-period pointers and flags.
template <typename BT>
struct R16Period {
  const BT *bvCur;   // boundary line + rst0 + 15 (reads of this period)
  const BT *bvNext;  // boundary line + rst0' + 15 (first group of the next period)
  BT *w;             // boundary line + rst0 (writes of the group's last lane)
  BT *wm;            // boundary line + Np + rst0 (mirror writes; reset periods only)
  bool ones;             // Special: the values written are ones
  int sCap;              // Special: step of the result cell in this period, or -1
  int rsi;               // Special: step at which this lane resets (its rho), or -1
};

// Group I (4 steps) of a period with ring parity PAR (slots 16 PAR + s). Each group loads the
// next group's inputs; group 3 loads the next period's group 0 (ring parity 1 - PAR), after the
// producer has written chunk per+1 (between groups 2 and 3, once the reads of group 3 of this
// period were issued).
template <Arith A, int R, int PAR, class F, int I>
__device__ __forceinline__ void r16_group(const char *rb, const R16Period<typename R16Bnd<A>::T> &pp,
                                          double s, R16State<A, R> &st, R16In<A> &nxt,
                                          double &result) {
  using BT = typename R16Bnd<A>::T;
  const int lane = threadIdx.x & 63;
  const int rho = lane & 15;
  R16In<A> cur = nxt;
  if constexpr (I < 3)
    r16_load<A>(rb, 16 * PAR + 4 * (I + 1), pp.bvCur + 4 * (I + 1), nxt);
  else
    r16_load<A>(rb, 16 * (1 - PAR), pp.bvNext, nxt);
  BT vals[4];
  if constexpr (A == Arith::MixedF) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int sidx = 4 * I + k;
      float a, b;
      coefs_f(cur.c[k], st.h, st.q, a, b);
      const float up = shift_down_ff<16>(st.k[R - 1], cur.bv[k]);
      bool z = false;
      if constexpr (F::kReset) z = pp.rsi == sidx;
      float above = up, aboveOld = st.kPrev;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float old = st.k[r];
        float kn = cell_f(above, old, aboveOld, a, b);
        if constexpr (F::kReset) kn = z ? 1.0f : kn;
        st.k[r] = kn;
        aboveOld = old;
        above = kn;
      }
      st.kPrev = up;
      if constexpr (F::kReset) {
        st.h = z ? st.hN : st.h;
        st.q = z ? st.qN : st.q;
      }
      if constexpr (F::kCapture) {
        if (sidx == pp.sCap) result = st.k[R - 1];
      }
      vals[k] = st.k[R - 1];
    }
  } else {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int sidx = 4 * I + k;
    double cA, cB;
    coefs<A>(cur.c[k], s, st.h, st.q, cA, cB);
    double up;
    if constexpr (ArithTraits<A>::kRound)  // float cells and float lines
      up = shift_down_f<16>(st.k[R - 1], cur.bv[k]);
    else
      up = shift_down<16>(st.k[R - 1], (double)cur.bv[k]);
    bool z = false;
    if constexpr (F::kReset) z = pp.rsi == sidx;
    double above = up, aboveOld = st.kPrev;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double old = st.k[r];
      double kn = cell<A>(above, old, aboveOld, cA, cB);
      if constexpr (F::kReset) kn = z ? 1.0 : kn;
      st.k[r] = kn;
      aboveOld = old;
      above = kn;
    }
    st.kPrev = up;
    if constexpr (F::kReset) {
      if constexpr (A == Arith::Mixed) {
        st.h = z ? st.hN : st.h;
        st.q = z ? st.qN : st.q;
      }
    }
    if constexpr (F::kCapture) {
      if (sidx == pp.sCap) result = st.k[R - 1];
    }
    vals[k] = (BT)st.k[R - 1];
  }
  }
  if (rho == 15) {
    BT *w = pp.w + 4 * I;
    if constexpr (F::kOnes) {
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = (BT)1;
    } else if constexpr (F::kReset) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const BT v = pp.ones ? (BT)1 : vals[k];
        w[k] = v;
        if (pp.wm) pp.wm[4 * I + k] = v;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = vals[k];
    }
  }
}

template <Arith A, int R, int PAR, class F>
__device__ __forceinline__ void r16_head(const char *rb, const R16Period<typename R16Bnd<A>::T> &pp,
                                         double s, R16State<A, R> &st, R16In<A> &nxt,
                                         double &result) {
  r16_group<A, R, PAR, F, 0>(rb, pp, s, st, nxt, result);
  r16_group<A, R, PAR, F, 1>(rb, pp, s, st, nxt, result);
  r16_group<A, R, PAR, F, 2>(rb, pp, s, st, nxt, result);
}

// Solves the wave's stream (g.stream pairs per lane group, one after another). For every pair of
// the stream, out(sp, value) is called by the lanes holding row M-1 (lane ((M-1) mod 16R) / R, its
// last row) with that group's k[M][N]. g.R must equal R.
//   P: producer (Global16Producer or the MFMA producer): prologue / early / period<PAR> /
//      lane_scale (Arith::Mixed: the scale of the pair whose chunk was multiplied last, read at
//      the start of the period that holds that pair's first reset)
//   s: per-lane double scale of the Scaled* modes (ignored otherwise)
//   lds: r16_lds_bytes<A>(N) bytes of this wave's LDS (GL: r16_ring_bytes<A>())
//   GL, gline: the four boundary lines in global memory at gline (r16_line_bytes<A>(N) bytes,
//      this wave's own) instead of LDS, for rows of columns longer than LDS holds. A line value is
//      read Np - 15 >= 17 steps after the same wave stored it; a wave's memory instructions reach
//      an address in program order, so the read sees the store.
template <Arith A, int R, bool GL = false, class Producer, class Out>
__device__ void ring16_solve(Producer &P, const PairGeom<16> &g, double s, void *lds, Out &&out,
                             typename R16Bnd<A>::T *gline = nullptr) {
  using RT = typename ArithTraits<A>::RT;
  using BT = typename R16Bnd<A>::T;
  const int lane = threadIdx.x & 63;
  const int rho = lane & 15;
  const int Np = g.Np;
  const int bs = r16_line_stride<A>(Np);
  BT *bnd;
  RT *ring;
  if constexpr (GL) {
    bnd = gline;
    ring = reinterpret_cast<RT *>(lds);
  } else {
    bnd = reinterpret_cast<BT *>(lds);
    ring = reinterpret_cast<RT *>(bnd + 4 * bs);
  }
  BT *line = bnd + (lane >> 4) * bs;
  const char *rb = reinterpret_cast<const char *>(ring + lane);

  for (int i = lane; i < kR16Slots * kR16Stride; i += 64) ring[i] = (RT)0;
  for (int i = lane; i < 4 * bs; i += 64) bnd[i] = (BT)1;

  const int total = g.columns();
  const int nperiods = total / kChunk + 1;  // the last lane finishes at step total + 14
  const int pc = g.pair_columns();
  const int bM = (g.M - 1) / (16 * R);
  const int rM = ((g.M - 1) - bM * 16 * R) / R;
  int tM = bM * Np + g.N + rM;  // result step of stream pair sp
  int sp = 0;

  P.prologue(ring, g);

  R16State<A, R> st;
#pragma unroll
  for (int r = 0; r < R; ++r) st.k[r] = 1.0;
  st.kPrev = 1.0;
  st.h = st.q = st.hN = st.qN = 0.0;
  R16In<A> nxt;  // inputs of the next group
  r16_load<A>(rb, 0, line + 15, nxt);

  int rst0 = 0;   // t0 mod Np
  int band = 0;   // band of column t0 within its pair
  for (int per = 0; per < nperiods; ++per) {
    const bool reset = rst0 == 0;
    const int rn = rst0 + kChunk == Np ? 0 : rst0 + kChunk;
    R16Period<BT> pp;
    pp.bvCur = line + rst0 + 15;
    pp.bvNext = line + rn + 15;
    pp.w = line + rst0;
    pp.wm = reset ? line + Np : nullptr;
    // the group's last lane is in the columns t0-15 .. t0; the band of t0-1 decides
    pp.ones = reset ? band == 0 : band == g.nb - 1;
    const bool cap = sp < g.stream && (tM >> 4) == per;
    pp.sCap = cap ? tM & 15 : -1;
    pp.rsi = reset ? rho : -1;
    if constexpr (A == Arith::Mixed || A == Arith::MixedF) {
      if (reset && band == 0) {
        const double m = P.lane_scale();
        st.hN = 0.5 * m;
        st.qN = m * m * (1.0 / 12);
      }
    }
    double result = 0.0;
    const bool odd = per & 1;
    const int fl = (reset || cap) ? 2 : (pp.ones ? 1 : 0);
    P.early(ring, g, per);
    // groups 0..2
    if (fl == 2) {
      if (odd) r16_head<A, R, 1, R16Special>(rb, pp, s, st, nxt, result);
      else r16_head<A, R, 0, R16Special>(rb, pp, s, st, nxt, result);
    } else if (fl == 1) {
      if (odd) r16_head<A, R, 1, R16Ones>(rb, pp, s, st, nxt, result);
      else r16_head<A, R, 0, R16Ones>(rb, pp, s, st, nxt, result);
    } else {
      if (odd) r16_head<A, R, 1, R16Plain>(rb, pp, s, st, nxt, result);
      else r16_head<A, R, 0, R16Plain>(rb, pp, s, st, nxt, result);
    }
    // chunk per+1 (parity of per+1) goes to the ring after the reads of group 3 were issued
    if (odd) P.template period<0>(ring, g, per);
    else P.template period<1>(ring, g, per);
    // group 3
    if (fl == 2) {
      if (odd) r16_group<A, R, 1, R16Special, 3>(rb, pp, s, st, nxt, result);
      else r16_group<A, R, 0, R16Special, 3>(rb, pp, s, st, nxt, result);
    } else if (fl == 1) {
      if (odd) r16_group<A, R, 1, R16Ones, 3>(rb, pp, s, st, nxt, result);
      else r16_group<A, R, 0, R16Ones, 3>(rb, pp, s, st, nxt, result);
    } else {
      if (odd) r16_group<A, R, 1, R16Plain, 3>(rb, pp, s, st, nxt, result);
      else r16_group<A, R, 0, R16Plain, 3>(rb, pp, s, st, nxt, result);
    }
    if (cap) {
      if (rho == rM) out(sp, result);
      ++sp;
      tM += pc;
    }
    rst0 = rn;
    if (rn == 0 && ++band == g.nb) band = 0;
  }
}

}  // namespace sigpde
