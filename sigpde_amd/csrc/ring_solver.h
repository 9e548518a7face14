// Goursat PDE wavefronts on one gfx950 wave (64 lanes), fed through an LDS ring.
//
// Scheme (stojl/SigPDE sigpde/cuda_device_functions.py:3-5, sigpde/cuda_pairwise_kernels.py:8-39):
//   k[i][0] = k[0][j] = 1;  k[i][j] = (k[i-1][j] + k[i][j-1]) A - k[i-1][j-1] B
//   A = 1 + c/2 + c^2/12,  B = 1 - c^2/12,  c = Delta[(i-1) >> order][(j-1) >> order] (* scale)
// Every cell reads only its three neighbours, so any dependency-respecting traversal reproduces the
// reference's anti-diagonal sweep value for value.
//
// Mapping (G = lanes per pair, 16 or 64; a wave solves P = 64/G pairs in lockstep)
//   * Lane rho of a group owns refined row G*b + rho of row band b. Bands run back to back as one
//     stream of "virtual columns" v = b*Np + jv, jv = 0 (a reset column) .. N, padded to Np.
//   * Lane rho handles virtual column v at step t = v + rho. The value above comes from lane
//     rho-1's previous step through a DPP shift (row_shr:1 inside 16-lane rows for G = 16,
//     wave_shr:1 for G = 64); the group's first lane takes the previous band's last row, which the
//     group's last lane parks in an LDS line `bnd` of Np values.
//   * At a reset column the lane's left value and corner become 1. Np is a multiple of the period
//     (kChunk steps), so resets happen in a few known periods; the other periods skip the select.
//   * Increments reach the lanes through an LDS ring of kRingW virtual columns x 64 rows (the P
//     pairs' rows side by side), filled one chunk of kChunk columns ahead by a Producer (coalesced
//     global loads here; the fused kernels produce with MFMA). The producer writes chunk p+1 just
//     before the last consumer group of period p, when the oldest live column is t0+1, so the ring
//     spans only the skew (G columns) plus one chunk. Lanes read the ring on their skewed
//     diagonal; kRingMirror duplicated slots let a lane read kU consecutive columns without
//     wrapping.
//   * Columns outside 1..N of a band (the reset column and the padding) may hold any increment:
//     the reset column's cell is replaced by 1 and padding cells only feed padding cells.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sigpde {

constexpr int kWave = 64;
constexpr int kChunk = 16;       // virtual columns per producer chunk (= steps per period)
constexpr int kU = 4;            // consumer group: columns read per LDS batch
constexpr int kRingMirror = 4;   // slots W..W+3 mirror slots 0..3
constexpr int kRingStride = 66;  // elements per slot: 64 rows + 2 pad; conflict-free skewed reads
                                 // (bank = 2t - rho + 16g) and chunk writes (bank = 2col + row)

template <int G>
struct Layout {
  static_assert(G == 16 || G == 64, "lanes per pair");
  static constexpr int kPairs = kWave / G;
  static constexpr int kRingW = G == 64 ? 80 : 32;     // >= G + kChunk, multiple of kChunk
  static constexpr int kRingSlots = kRingW + kRingMirror;
  static constexpr int kMinBand = G == 64 ? 80 : 32;   // > G + 2*kU: bnd read after its write
  static_assert(kRingW >= G + kChunk && kRingW % kChunk == 0 && G % kChunk == 0,
                "ring geometry");
  __host__ __device__ static int band_len(int N) {
    int np = N + 1 > kMinBand ? N + 1 : kMinBand;
    return (np + kChunk - 1) / kChunk * kChunk;
  }
};

// ---------------------------------------------------------------- cross-lane helpers
template <int G>
__device__ __forceinline__ int dpp_down1(int v, int old) {
  // lane rho receives lane rho-1 of its group; the group's first lane receives `old`
  if constexpr (G == 64)
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);  // wave_shr:1
  else
    return __builtin_amdgcn_update_dpp(old, v, 0x111, 0xf, 0xf, false);  // row_shr:1
}
template <int G>
__device__ __forceinline__ double shift_down(double v, double old) {
  int2 a = *reinterpret_cast<int2 *>(&v);
  int2 o = *reinterpret_cast<int2 *>(&old);
  int2 r;
  r.x = dpp_down1<G>(a.x, o.x);
  r.y = dpp_down1<G>(a.y, o.y);
  return *reinterpret_cast<double *>(&r);
}
// Float form of shift_down for the float32 modes (Arith with kRound), whose cell values are all
// floats: the float is shifted (one DPP move instead of two; `top` a float too), the same value.
template <int G>
__device__ __forceinline__ double shift_down_f(double v, float old) {
  return (double)__int_as_float(dpp_down1<G>(__float_as_int((float)v), __float_as_int(old)));
}
// The float32 wavefront's shift (Arith::MixedF): a float shifted as a float.
template <int G>
__device__ __forceinline__ float shift_down_ff(float v, float old) {
  return __int_as_float(dpp_down1<G>(__float_as_int(v), __float_as_int(old)));
}
__device__ __forceinline__ double readlane_d(double v, int l) {
  int2 a = *reinterpret_cast<int2 *>(&v);
  int2 r;
  r.x = __builtin_amdgcn_readlane(a.x, l);
  r.y = __builtin_amdgcn_readlane(a.y, l);
  return *reinterpret_cast<double *>(&r);
}

// ---------------------------------------------------------------- arithmetic modes
// Ref32: ring holds float c = float(inc * scale); c*c in float, the rest in double, and every cell
//        is rounded to float as it is stored (the reference's float32 typing).
// Ref64: ring holds double c; all double.
// Scaled32 / Scaled64: ring holds the raw increment; c = double(inc) * s with a per-lane double s
//        (the root-find's inc * xt**2), cells rounded to float for Scaled32.
// Mixed: ring holds the raw float increment product p of the split-precision increment Gram; the
//        pair's scale m (c = m * p) is folded into per-lane double coefficients h = m/2 and
//        q = m^2/12; double arithmetic and state.
// MixedF: as Mixed with a float32 wavefront (float coefficients, arithmetic and state; 16-lane
//        groups only, ring16.h). The cell is written with the increment terms apart from the
//        ones, k = k01 + (k01 a + ((k10 - k00) + k10 a + k00 b)), a = c/2 + c^2/12, b = c^2/12:
//        c/2 and c^2/12 keep float precision instead of being rounded against 1 (which loses
//        c^2/12 entirely at dyadic order 2 and more), and k10 - k00 of neighbouring cells is
//        nearly exact.
enum class Arith { Ref32, Ref64, Scaled32, Scaled64, Mixed, MixedF };

// RT: ring element. VT: the wavefront's cell values and coefficients.
template <Arith A>
struct ArithTraits;
template <>
struct ArithTraits<Arith::Ref32> {
  using RT = float;
  using VT = double;
  static constexpr bool kRound = true;
};
template <>
struct ArithTraits<Arith::Ref64> {
  using RT = double;
  using VT = double;
  static constexpr bool kRound = false;
};
template <>
struct ArithTraits<Arith::Scaled32> {
  using RT = float;
  using VT = double;
  static constexpr bool kRound = true;
};
template <>
struct ArithTraits<Arith::Scaled64> {
  using RT = double;
  using VT = double;
  static constexpr bool kRound = false;
};
template <>
struct ArithTraits<Arith::Mixed> {
  using RT = float;
  using VT = double;
  static constexpr bool kRound = false;
};
template <>
struct ArithTraits<Arith::MixedF> {
  using RT = float;
  using VT = float;
  static constexpr bool kRound = false;
};

// MixedF: coefficients of the float32 wavefront from the raw product p (c = m p, h = m/2,
// q = m^2/12): b = q p^2, a = h p + b.
__device__ __forceinline__ void coefs_f(float p, float h, float q, float &a, float &b) {
  b = q * (p * p);
  a = __builtin_fmaf(h, p, b);
}
// MixedF cell: k = k01 + (k01 a + w), w = (k10 - k00) + k10 a + k00 b (w does not wait for the
// value shifted in from the row above).
__device__ __forceinline__ float cell_f(float k01, float k10, float k00, float a, float b) {
  const float w = __builtin_fmaf(k00, b, __builtin_fmaf(k10, a, k10 - k00));
  return __builtin_fmaf(k01, a, w) + k01;
}

// A = (1 + c/2) + c^2/12 and B = 1 - c^2/12 as pde_step writes them, with the products contracted
// into FMAs as NVVM contracts the reference's expression (fma(c, 0.5, 1) equals 1 + 0.5 c: the
// product is exact). The contractions are written out so that every kernel instantiation rounds
// a cell the same way: a pair's result does not depend on which launch shape solved it.
template <Arith A>
__device__ __forceinline__ void coefs(typename ArithTraits<A>::RT c, double s, double h, double q,
                                      double &cA, double &cB) {
  double cd, c2;
  if constexpr (A == Arith::Mixed) {
    // A = 1 + (m/2) p + (m^2/12) p^2, B = 1 - (m^2/12) p^2 with c = m p
    cd = (double)c;
    c2 = cd * cd;
    cA = fma(q, c2, fma(cd, h, 1.0));
    cB = fma(-q, c2, 1.0);
    return;
  } else if constexpr (A == Arith::Ref32) {
    float c2f = c * c;
    cd = (double)c;
    c2 = (double)c2f;
  } else if constexpr (A == Arith::Scaled32 || A == Arith::Scaled64) {
    cd = (double)c * s;
    c2 = cd * cd;
  } else {
    cd = (double)c;
    c2 = cd * cd;
  }
  cA = fma(c2, 1.0 / 12, fma(cd, 0.5, 1.0));
  cB = fma(c2, -1.0 / 12, 1.0);
}

// k = (k01 + k10) A - k00 B.
// The reference-scheme modes (Ref*, Scaled*) contract it as NVVM contracts pde_step on the
// reference's target, fma(k01 + k10, A, -(k00 B)), so that a cell rounds exactly as in the
// oracle's contracted setting. Mixed (whose increments come from MFMA products and are held only
// to a tolerance) groups it as k01*A + (k10*A - k00*B), so that only the last FMA waits for the
// value shifted in from the row above.
template <Arith A>
__device__ __forceinline__ double cell(double k01, double k10, double k00, double cA, double cB) {
  double r;
  if constexpr (A == Arith::Mixed) {
    const double w = fma(k10, cA, -(k00 * cB));
    r = fma(k01, cA, w);
  } else {
    r = fma(k01 + k10, cA, -(k00 * cB));
  }
  if constexpr (ArithTraits<A>::kRound) r = (double)(float)r;
  return r;
}

// LDS for one wave: P band-boundary lines of Np doubles, then the ring (the lines come first so
// that their per-group offsets fit the LDS instructions' immediate fields).
template <int G>
__host__ __device__ inline int bnd_stride(int N) { return Layout<G>::band_len(N); }
template <Arith A, int G>
__host__ __device__ inline size_t wave_lds_bytes(int N) {
  using L = Layout<G>;
  return (size_t)L::kRingSlots * kRingStride * sizeof(typename ArithTraits<A>::RT) +
         (size_t)L::kPairs * bnd_stride<G>(N) * sizeof(double);
}

// Geometry of a wave's virtual-column stream: `stream` consecutive pairs (all P lane groups of a
// wave, and all pairs of the stream, share it), each pair nb bands of Np columns.
template <int G>
struct PairGeom {
  int m, n, order;  // original increments and dyadic order
  int M, N;         // refined rows / columns
  int Np;           // virtual columns per band
  int R;            // refined rows per lane (a power of two dividing 2^order; ring16.h)
  int nb;           // bands of G * R rows
  int stream;       // pairs solved one after another
  __device__ __host__ PairGeom(int m_, int n_, int o_, int stream_ = 1, int R_ = 1)
      : m(m_), n(n_), order(o_), R(R_), stream(stream_) {
    M = m << o_;
    N = n << o_;
    Np = Layout<G>::band_len(N);
    nb = (M + G * R - 1) / (G * R);
  }
  __device__ __host__ int pair_columns() const { return nb * Np; }
  __device__ __host__ int columns() const { return stream * nb * Np; }
  // chunk k -> (stream pair, band, band column of the chunk's first virtual column)
  __device__ __forceinline__ void locate(int k, int &sp, int &band, int &jv0) const {
    const int v0 = k * kChunk;
    const int pc = nb * Np;
    sp = v0 / pc;
    const int r = v0 - sp * pc;
    band = r / Np;
    jv0 = r - band * Np;
  }
};

// Position of a chunk in the stream, advanced one chunk at a time (no divisions on the hot path).
template <int G>
struct Cursor {
  int sp, band, jv0;  // stream pair, band, band column of the chunk's first virtual column
  __device__ __forceinline__ void start(const PairGeom<G> &g, int k) { g.locate(k, sp, band, jv0); }
  __device__ __forceinline__ void advance(const PairGeom<G> &g) {
    jv0 += kChunk;
    if (jv0 == g.Np) {
      jv0 = 0;
      if (++band == g.nb) {
        band = 0;
        ++sp;
      }
    }
  }
};

// Producer interface (used by ring_solve):
//   prologue(ring, geom): write chunk 0 and start chunk 1;
//   early(ring, geom, p): called at the start of period p (before the consumer's first group);
//   period(ring, geom, p): called before the last consumer group of period p; writes chunk p+1.
//
// GlobalProducer reads increments (m x n, row stride ld) of type RT from global memory; lane group
// g serves the pair whose increments start at inc[g] (null: an idle slot, served zeros). The
// stream holds one pair per group.
// Chunk k covers virtual columns [k*kChunk, (k+1)*kChunk) of one band (Np is a multiple of
// kChunk): 64 ring rows (P pairs x G refined rows) x 16 columns. Load i (of 16) has lane l fetch
// ring row 4i + (l >> 4) at column l & 15, i.e. four 64-byte row segments per load. load() only
// issues the loads (indices clamped into the increment array); scaling happens in store(), one
// period later, so the loads stay in flight while the wave computes. Rows past M and columns
// outside 1..N read clamped data, which never reaches a result cell.
template <typename RT, int G>
struct GlobalProducer {
  using L = Layout<G>;
  const RT *inc[L::kPairs];
  RT scale[L::kPairs];
  int64_t ld;
  bool scaled;
  RT buf[kChunk];

  Cursor<G> cl;  // next chunk to load
  int slotS;     // ring slot of the next chunk to store

  __device__ __forceinline__ void prologue(RT *ring, const PairGeom<G> &g) {
    cl.start(g, 0);
    slotS = 0;
    load(g);
    store(ring);
    if (g.columns() > kChunk) load(g);
  }
  __device__ __forceinline__ void early(RT *, const PairGeom<G> &, int) {}
  __device__ __forceinline__ void period(RT *ring, const PairGeom<G> &g, int per) {
    const int nchunks = g.columns() / kChunk;
    if (per + 1 < nchunks) store(ring);
    if (per + 2 < nchunks) load(g);
  }

  __device__ __forceinline__ void load(const PairGeom<G> &g) {
    const int lane = threadIdx.x & (kWave - 1);
    const int b = cl.band, jv0 = cl.jv0;
    const int c = lane & 15;
    int jo = (jv0 + c - 1) >> g.order;
    jo = jo < 0 ? 0 : (jo > g.n - 1 ? g.n - 1 : jo);
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int rr = 4 * i + (lane >> 4);  // ring row
      const int pg = rr / G;               // pair slot
      const int rho = rr - pg * G;
      int ro = (b * G + rho) >> g.order;
      ro = ro > g.m - 1 ? g.m - 1 : ro;
      const RT *base = inc[pg];
      buf[i] = base ? base[(int64_t)ro * ld + jo] : (RT)0;
    }
    cl.advance(g);
  }
  __device__ __forceinline__ void store(RT *ring) {
    const int lane = threadIdx.x & (kWave - 1);
    const int c = lane & 15;
    const int slot0 = slotS;
    RT v[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int pg = (4 * i + (lane >> 4)) / G;
      v[i] = scaled ? buf[i] * scale[pg] : buf[i];
    }
    RT *dst = ring + (slot0 + c) * kRingStride + (lane >> 4);
#pragma unroll
    for (int i = 0; i < kChunk; ++i) dst[4 * i] = v[i];
    if (slot0 == 0 && c < kRingMirror) {
      RT *mir = ring + (L::kRingW + c) * kRingStride + (lane >> 4);
#pragma unroll
      for (int i = 0; i < kChunk; ++i) mir[4 * i] = v[i];
    }
    slotS += kChunk;
    if (slotS == L::kRingW) slotS = 0;
  }
};

// Ring and boundary values of one consumer group (kU steps), loaded one group ahead.
template <Arith A>
struct GroupIn {
  typename ArithTraits<A>::RT cv[kU];
  double bv[kU];
};

// roff: byte offset (from the wave's LDS base) of this lane's ring element at the group's first
// column; bp: the group's boundary values.
template <Arith A>
__device__ __forceinline__ void load_group(const char *lds, uint32_t roff, const double *bp,
                                           GroupIn<A> &in) {
  using RT = typename ArithTraits<A>::RT;
  const RT *rp = reinterpret_cast<const RT *>(lds + roff);
#pragma unroll
  for (int q = 0; q < kU; ++q) in.cv[q] = rp[q * kRingStride];
#pragma unroll
  for (int q = 0; q < kU; ++q) in.bv[q] = bp[q];
}

// Consumer state carried from group to group.
template <Arith A>
struct ConsumerState {
  GroupIn<A> nxt;   // inputs of the next group
  uint32_t roff;    // ring byte offset of the group held in nxt
  uint32_t rlim;    // roff wraps (by kRingW slots) when it reaches this
  double k, kPrev;
  double carry;     // the group's last lane: value of the previous group's last step
  double h, q;      // Mixed: coefficients of the lane's current pair
  double hN, qN;    // Mixed: coefficients taken at the lane's next reset
};

// Per-period boundary-line write plan: the group's last lane writes, one step late, the block of kU
// columns starting at wp + kU*i for group i (at the group starting at t0 it holds columns
// t0-(G-1)+q, q < kU, i.e. the block [t0-G, t0-G+kU) = {previous group's last value, this group's
// first kU-1}). ones: the blocks belong to a pair's last band; they are written as ones, the next
// pair's top boundary.
struct WritePlan {
  double *wp;
  bool en;
  bool ones;
};

// Consumer-period flavours: kReset (a lane may reach a reset column), kCapture (the result cell
// may fall in the period), kOnes (the write plan may ask for ones). Plain periods have none.
template <bool R, bool C, bool O>
struct Flavour {
  static constexpr bool kReset = R, kCapture = C, kOnes = O;
};
using Plain = Flavour<false, false, false>;
using ResetOnly = Flavour<true, false, false>;
using Special = Flavour<true, true, true>;

// Group i (of 4) of a period: kU steps starting at t = t0 + 4i.
//   bnext: boundary values the group prefetches for the next group; dl: rho - rst0 - 4i (the lane
//   resets at step t + dl); tM, rM: result step and lane.
template <Arith A, int G, class F>
__device__ __forceinline__ void consume_group(const char *lds, const double *bnext, double s,
                                              int t, int dl, double *wblk, const WritePlan &w,
                                              GroupIn<A> &cur, ConsumerState<A> &st, int tM,
                                              int rM, double &result) {
  constexpr uint32_t kRingBytes = Layout<G>::kRingW * kRingStride * sizeof(typename ArithTraits<A>::RT);
  constexpr uint32_t kGroupBytes = kU * kRingStride * sizeof(typename ArithTraits<A>::RT);
  const int lane = threadIdx.x & (kWave - 1);
  const int rho = lane & (G - 1);
  // prefetch the next group's inputs
  st.roff += kGroupBytes;
  st.roff = st.roff >= st.rlim ? st.roff - kRingBytes : st.roff;
  load_group<A>(lds, st.roff, bnext, st.nxt);
  double last[kU];
#pragma unroll
  for (int q = 0; q < kU; ++q) {
    double cA, cB;
    coefs<A>(cur.cv[q], s, st.h, st.q, cA, cB);
    const double up = shift_down<G>(st.k, cur.bv[q]);
    const double kn = cell<A>(up, st.k, st.kPrev, cA, cB);
    st.kPrev = up;
    st.k = kn;
    if constexpr (F::kReset) {
      const bool z = dl == q;
      st.k = z ? 1.0 : st.k;
      st.kPrev = z ? 1.0 : st.kPrev;
      if constexpr (A == Arith::Mixed) {
        st.h = z ? st.hN : st.h;
        st.q = z ? st.qN : st.q;
      }
    }
    if constexpr (F::kCapture) {
      if (t + q == tM) result = rho == rM ? st.k : result;
    }
    last[q] = st.k;
  }
  if (w.en && rho == G - 1) {
    if (F::kOnes && w.ones) {
      wblk[0] = 1.0;
      wblk[1] = 1.0;
      wblk[2] = 1.0;
      wblk[3] = 1.0;
    } else {
      wblk[0] = st.carry;
      wblk[1] = last[0];
      wblk[2] = last[1];
      wblk[3] = last[2];
    }
  }
  st.carry = last[kU - 1];
}

// Groups 0..2 of a period (rst0 = t0 mod Np: the group's first lane's band column, and the lane
// that resets at t0). Group i prefetches boundary values at bndLine + rst0 + kU*(i+1).
template <Arith A, int G, class F>
__device__ __forceinline__ void consume_head(const char *lds, const double *bndLine, double s,
                                             int t0, int rst0, const WritePlan &w,
                                             ConsumerState<A> &st, int tM, int rM,
                                             double &result) {
  const int lane = threadIdx.x & (kWave - 1);
  const int dl = (lane & (G - 1)) - rst0;
  const double *bp = bndLine + rst0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    GroupIn<A> cur = st.nxt;
    consume_group<A, G, F>(lds, bp + kU * (i + 1), s, t0 + kU * i, dl - kU * i, w.wp + kU * i, w,
                           cur, st, tM, rM, result);
  }
}

// Group 3 of a period; it prefetches the next period's first boundary block at bndLine + jnext.
template <Arith A, int G, class F>
__device__ __forceinline__ void consume_tail(const char *lds, const double *bndLine, double s,
                                             int t0, int rst0, int jnext, const WritePlan &w,
                                             ConsumerState<A> &st, int tM, int rM,
                                             double &result) {
  const int lane = threadIdx.x & (kWave - 1);
  const int dl = (lane & (G - 1)) - rst0;
  GroupIn<A> cur = st.nxt;
  consume_group<A, G, F>(lds, bndLine + jnext, s, t0 + 3 * kU, dl - 3 * kU, w.wp + 3 * kU, w, cur,
                         st, tM, rM, result);
}

// Solves the wave's stream: g.stream pairs per lane group, one after another. For every pair of
// the stream, out(sp, value) is called by the lanes rho == (M-1) mod G of each group with that
// group's k[M][N] (sp = position in the stream).
//   P: producer (interface above); for Arith::Mixed it also provides lane_scale(), the scale m of
//      the pair whose chunk it multiplied last (read at the start of the period that holds a
//      pair's first reset, before early(): the last chunk multiplied then is that pair's first)
//   s: the per-lane double scale of the Scaled* modes (ignored otherwise)
//   lds: this wave's LDS (wave_lds_bytes<A, G>(N) bytes)
template <Arith A, int G, class Producer, class Out>
__device__ void ring_solve(Producer &P, const PairGeom<G> &g, double s, void *lds, Out &&out) {
  using RT = typename ArithTraits<A>::RT;
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int rho = lane & (G - 1);
  const int bs = bnd_stride<G>(g.N);
  double *bnd = reinterpret_cast<double *>(lds);
  RT *ring = reinterpret_cast<RT *>(bnd + L::kPairs * bs);
  const char *ldsb = reinterpret_cast<const char *>(lds);
  const uint32_t ringOff = (uint32_t)(L::kPairs * bs * sizeof(double));
  double *bndLine = bnd + (lane / G) * bs;  // this lane's pair's boundary line

  // zero the ring (columns not yet produced read as c = 0) and set the boundary lines to 1
  for (int i = lane; i < L::kRingSlots * kRingStride; i += kWave) ring[i] = (RT)0;
  for (int i = lane; i < L::kPairs * bs; i += kWave) bnd[i] = 1.0;

  const int total = g.columns();            // virtual columns
  const int nsteps = total + G - 1;         // last lane finishes here
  const int nperiods = (nsteps + kChunk - 1) / kChunk;
  const int pc = g.pair_columns();

  // result cell (M, N) of stream pair sp: band bM, lane rM of each group, step tM + sp * pc
  const int bM = (g.M - 1) / G;
  const int rM = (g.M - 1) - bM * G;
  int tM = bM * g.Np + g.N + rM;
  int sp = 0;

  P.prologue(ring, g);

  ConsumerState<A> st;
  st.k = 1.0;
  st.kPrev = 1.0;
  st.carry = 1.0;
  st.h = st.q = st.hN = st.qN = 0.0;
  static_assert(A != Arith::Mixed || G == kChunk, "Mixed: one reset period per band");
  int rband = 0;  // band whose reset column the next reset period holds
  st.rlim = ringOff + (uint32_t)((L::kRingW * kRingStride + lane) * sizeof(RT));
  // ring slot at step 0 is virtual column -rho
  st.roff = ringOff + (uint32_t)((((L::kRingW - rho) % L::kRingW) * kRingStride + lane) * sizeof(RT));
  load_group<A>(ldsb, st.roff, bndLine, st.nxt);
  int rst0 = 0;  // t0 mod Np: the group's first lane's band column
  // The group starting at step t writes the boundary block of virtual column t - G (nothing while
  // that is negative). G is a multiple of kChunk, so a period writes either nothing or four blocks
  // of one band, at band column wjv of band wband (tracked without divisions).
  int vw = -G;
  int wjv = 0, wband = 0;

  for (int per = 0; per < nperiods; ++per) {
    const int t0 = per * kChunk;
    WritePlan w;
    w.en = vw >= 0;
    w.wp = bndLine + wjv;
    w.ones = w.en && wband == g.nb - 1;
    if (w.en) {
      wjv += kChunk;
      if (wjv == g.Np) {
        wjv = 0;
        if (++wband == g.nb) wband = 0;
      }
    }
    int jnext = rst0 + kChunk;
    jnext = jnext == g.Np ? 0 : jnext;
    if constexpr (A == Arith::Mixed) {
      if (rst0 == 0) {
        if (rband == 0) {
          const double m = P.lane_scale();
          st.hN = 0.5 * m;
          st.qN = m * m * (1.0 / 12);
        }
        if (++rband == g.nb) rband = 0;
      }
    }
    P.early(ring, g, per);
    double result = 0.0;
    const bool cap = sp < g.stream && tM >= t0 && tM < t0 + kChunk;
    const int fl = (cap || w.ones) ? 2 : (rst0 < G ? 1 : 0);
    if (fl == 2)
      consume_head<A, G, Special>(ldsb, bndLine, s, t0, rst0, w, st, tM, rM, result);
    else if (fl == 1)
      consume_head<A, G, ResetOnly>(ldsb, bndLine, s, t0, rst0, w, st, tM, rM, result);
    else
      consume_head<A, G, Plain>(ldsb, bndLine, s, t0, rst0, w, st, tM, rM, result);
    P.period(ring, g, per);
    if (fl == 2) {
      consume_tail<A, G, Special>(ldsb, bndLine, s, t0, rst0, jnext, w, st, tM, rM, result);
      if (cap) {
        if (rho == rM) out(sp, result);
        ++sp;
        tM += pc;
      }
    } else if (fl == 1) {
      consume_tail<A, G, ResetOnly>(ldsb, bndLine, s, t0, rst0, jnext, w, st, tM, rM, result);
    } else {
      consume_tail<A, G, Plain>(ldsb, bndLine, s, t0, rst0, jnext, w, st, tM, rM, result);
    }
    rst0 += kChunk;
    if (rst0 == g.Np) rst0 = 0;
    vw += kChunk;
  }
}

}  // namespace sigpde
