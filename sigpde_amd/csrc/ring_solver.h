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
//     global loads here; the fused kernels produce with MFMA). Lanes read the ring on their skewed
//     diagonal; kRingMirror duplicated slots let a lane read kU consecutive columns without
//     wrapping. The ring spans the skew (G columns) plus two chunks, so G = 16 needs half the LDS of
//     G = 64 and lets twice as many waves share a CU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sigpde {

constexpr int kWave = 64;
constexpr int kChunk = 16;       // virtual columns per producer chunk (= steps per period)
constexpr int kU = 4;            // consumer group: columns read per LDS batch
constexpr int kRingMirror = 8;   // slots W..W+7 mirror slots 0..7
constexpr int kRingStride = 66;  // elements per slot: 64 rows + 2 pad; conflict-free skewed reads
                                 // (bank = 2t - rho + 16g) and chunk writes (bank = 2col + row)

template <int G>
struct Layout {
  static_assert(G == 16 || G == 64, "lanes per pair");
  static constexpr int kPairs = kWave / G;
  static constexpr int kRingW = G == 64 ? 96 : 48;     // >= G + 2*kChunk, multiple of kChunk
  static constexpr int kRingSlots = kRingW + kRingMirror;
  static constexpr int kMinBand = G == 64 ? 80 : 32;   // > G + 2*kU: bnd read after its write
  static_assert(kRingW >= G + 2 * kChunk && kRingW % kChunk == 0, "ring geometry");
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
// Mixed: ring holds float c from the split-precision increment Gram; double arithmetic and state.
enum class Arith { Ref32, Ref64, Scaled32, Scaled64, Mixed };

template <Arith A>
struct ArithTraits;
template <>
struct ArithTraits<Arith::Ref32> { using RT = float; static constexpr bool kRound = true; };
template <>
struct ArithTraits<Arith::Ref64> { using RT = double; static constexpr bool kRound = false; };
template <>
struct ArithTraits<Arith::Scaled32> { using RT = float; static constexpr bool kRound = true; };
template <>
struct ArithTraits<Arith::Scaled64> { using RT = double; static constexpr bool kRound = false; };
template <>
struct ArithTraits<Arith::Mixed> { using RT = float; static constexpr bool kRound = false; };

// A = (1 + c/2) + c^2/12 and B = 1 - c^2/12 as pde_step writes them (the compiler may contract the
// products into FMAs, as NVVM does for the reference).
template <Arith A>
__device__ __forceinline__ void coefs(typename ArithTraits<A>::RT c, double s, double &cA,
                                      double &cB) {
  double cd, c2;
  if constexpr (A == Arith::Ref32) {
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
  cA = (1.0 + 0.5 * cd) + (1.0 / 12) * c2;
  cB = 1.0 - (1.0 / 12) * c2;
}

// k = (k01 + k10) A - k00 B, grouped as k01*A + (k10*A - k00*B) so that only the last FMA waits
// for the value shifted in from the row above.
template <Arith A>
__device__ __forceinline__ double cell(double k01, double k10, double k00, double cA, double cB) {
  double w = k10 * cA - k00 * cB;
  double r = k01 * cA + w;
  if constexpr (ArithTraits<A>::kRound) r = (double)(float)r;
  return r;
}

// LDS for one wave: the ring, then P band-boundary lines of Np doubles.
template <Arith A, int G>
__host__ __device__ inline size_t wave_lds_bytes(int N) {
  using L = Layout<G>;
  return (size_t)L::kRingSlots * kRingStride * sizeof(typename ArithTraits<A>::RT) +
         (size_t)L::kPairs * L::band_len(N) * sizeof(double);
}

// Geometry of a wave's virtual-column stream: `stream` consecutive pairs (all P lane groups of a
// wave, and all pairs of the stream, share it), each pair nb bands of Np columns.
template <int G>
struct PairGeom {
  int m, n, order;  // original increments and dyadic order
  int M, N;         // refined rows / columns
  int Np;           // virtual columns per band
  int nb;           // bands of G rows
  int stream;       // pairs solved one after another
  __device__ __host__ PairGeom(int m_, int n_, int o_, int stream_ = 1)
      : m(m_), n(n_), order(o_), stream(stream_) {
    M = m << o_;
    N = n << o_;
    Np = Layout<G>::band_len(N);
    nb = (M + G - 1) / G;
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

// Producer that reads increments (m x n, row stride ld) of type RT from global memory; lane group
// g serves the pair whose increments start at inc[g] (null: an idle slot, served zeros). The
// stream holds one pair per group.
// Chunk k covers virtual columns [k*kChunk, (k+1)*kChunk) of one band (Np is a multiple of
// kChunk): 64 ring rows (P pairs x G refined rows) x 16 columns. Load i (of 16) has lane l fetch
// ring row 4i + (l >> 4) at column l & 15, i.e. four 64-byte row segments per load. load() only
// issues the loads (clamped, always in bounds); masking and scaling happen in store(), one period
// later, so the loads stay in flight while the wave computes. Rows past M read clamped data: they
// only feed rows further down, never a row of the result.
template <typename RT, int G>
struct GlobalProducer {
  using L = Layout<G>;
  const RT *inc[L::kPairs];
  RT scale[L::kPairs];
  int64_t ld;
  bool scaled;
  RT buf[kChunk];
  int pendJv0;

  __device__ __forceinline__ void prologue(RT *ring, const PairGeom<G> &g) {
    load(g, 0);
    store(ring, g, 0);
    if (g.columns() > kChunk) load(g, 1);
  }
  __device__ __forceinline__ void begin_period(RT *ring, const PairGeom<G> &g, int per) {
    const int nchunks = g.columns() / kChunk;
    if (per + 1 < nchunks) store(ring, g, per + 1);
    if (per + 2 < nchunks) load(g, per + 2);
  }

  __device__ __forceinline__ void load(const PairGeom<G> &g, int k) {
    const int lane = threadIdx.x & (kWave - 1);
    int sp, b, jv0;
    g.locate(k, sp, b, jv0);
    const int c = lane & 15;
    int jo = (jv0 + c - 1) >> g.order;
    jo = jo < 0 ? 0 : (jo > g.n - 1 ? g.n - 1 : jo);
    const int bb = b < g.nb ? b : g.nb - 1;
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int rr = 4 * i + (lane >> 4);  // ring row
      const int pg = rr / G;               // pair slot
      const int rho = rr - pg * G;
      int ro = (bb * G + rho) >> g.order;
      ro = ro > g.m - 1 ? g.m - 1 : ro;
      const RT *base = inc[pg];
      buf[i] = base ? __builtin_nontemporal_load(base + (int64_t)ro * ld + jo) : (RT)0;
    }
    pendJv0 = jv0;
  }
  __device__ __forceinline__ void store(RT *ring, const PairGeom<G> &g, int k) {
    const int lane = threadIdx.x & (kWave - 1);
    const int c = lane & 15;
    const int slot0 = (k * kChunk) % L::kRingW;
    const int jv = pendJv0 + c;
    const bool ok = jv >= 1 && jv <= g.N;
    RT v[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int pg = (4 * i + (lane >> 4)) / G;
      RT x = scaled ? buf[i] * scale[pg] : buf[i];
      v[i] = ok ? x : (RT)0;
    }
    RT *dst = ring + (slot0 + c) * kRingStride + (lane >> 4);
#pragma unroll
    for (int i = 0; i < kChunk; ++i) dst[4 * i] = v[i];
    if (slot0 == 0 && c < kRingMirror) {
      RT *mir = ring + (L::kRingW + c) * kRingStride + (lane >> 4);
#pragma unroll
      for (int i = 0; i < kChunk; ++i) mir[4 * i] = v[i];
    }
  }
};

// One period (kChunk steps) of the consumer. kReset: lanes reach a reset column in this period;
// kCapture: the result cell is computed in this period.
template <Arith A, int G, bool kReset, bool kCapture>
__device__ __forceinline__ void consume_period(const typename ArithTraits<A>::RT *ring,
                                               double *bnd, const PairGeom<G> &g, double s,
                                               int t0, int rst0, int jw0, bool top, int &slot,
                                               double &k, double &kPrev, int tM, int rM,
                                               double &result) {
  using RT = typename ArithTraits<A>::RT;
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int rho = lane & (G - 1);
  const int grpBase = (lane / G) * g.Np;  // this lane's pair's bnd line
  const int dl = rho - rst0;              // this lane resets at step t0 + dl
#pragma unroll
  for (int grp = 0; grp < kChunk; grp += kU) {
    RT cv[kU];
    const RT *rp = ring + slot * kRingStride + lane;
#pragma unroll
    for (int q = 0; q < kU; ++q) cv[q] = rp[q * kRingStride];
    slot += kU;
    if (slot >= L::kRingW) slot -= L::kRingW;
    double bv[kU];
#pragma unroll
    for (int q = 0; q < kU; ++q) bv[q] = top ? 1.0 : bnd[grpBase + rst0 + grp + q];
    double last[kU];
#pragma unroll
    for (int q = 0; q < kU; ++q) {
      double cA, cB;
      coefs<A>(cv[q], s, cA, cB);
      const double up = shift_down<G>(k, bv[q]);
      const double kn = cell<A>(up, k, kPrev, cA, cB);
      kPrev = up;
      k = kn;
      if constexpr (kReset) {
        const bool z = dl == grp + q;
        k = z ? 1.0 : k;
        kPrev = z ? 1.0 : kPrev;
      }
      if constexpr (kCapture) {
        if (t0 + grp + q == tM) result = rho == rM ? k : result;
      }
      last[q] = k;
    }
    // the group's last lane parks its values (virtual columns t - (G-1)) for the next band
    if (rho == G - 1) {
      int j = jw0 + grp;
#pragma unroll
      for (int q = 0; q < kU; ++q) {
        int jj = j + q;
        jj = jj >= g.Np ? jj - g.Np : jj;
        bnd[grpBase + jj] = last[q];
      }
    }
  }
}

// Solves the wave's stream: g.stream pairs per lane group, one after another. For every pair of
// the stream, out(sp, value) is called by the lanes rho == (M-1) mod G of each group with that
// group's k[M][N] (sp = position in the stream).
//   P: producer with prologue(ring, geom) / begin_period(ring, geom, period)
//   s: the per-lane double scale of the Scaled* modes (ignored otherwise)
//   lds: this wave's LDS (wave_lds_bytes<A, G>(N) bytes)
template <Arith A, int G, class Producer, class Out>
__device__ void ring_solve(Producer &P, const PairGeom<G> &g, double s, void *lds, Out &&out) {
  using RT = typename ArithTraits<A>::RT;
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int rho = lane & (G - 1);
  RT *ring = reinterpret_cast<RT *>(lds);
  double *bnd = reinterpret_cast<double *>(ring + L::kRingSlots * kRingStride);

  // zero the ring (columns not yet produced read as c = 0) and set the boundary lines to 1
  for (int i = lane; i < L::kRingSlots * kRingStride; i += kWave) ring[i] = (RT)0;
  for (int i = lane; i < L::kPairs * g.Np; i += kWave) bnd[i] = 1.0;

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

  double k = 1.0, kPrev = 1.0;
  int slot = (L::kRingW - rho) % L::kRingW;  // ring slot at step 0 (virtual column -rho)
  int rst0 = 0;                              // t0 mod Np: the group's first lane's band column
  int jw0 = g.Np - (G - 1);                  // (t0 - (G-1)) mod Np: the group's last lane's
  int band0 = 0;                             // band of the group's first lane

  for (int per = 0; per < nperiods; ++per) {
    P.begin_period(ring, g, per);
    const int t0 = per * kChunk;
    const bool top = band0 == 0;
    double result = 0.0;
    if (sp < g.stream && tM >= t0 && tM < t0 + kChunk) {
      consume_period<A, G, true, true>(ring, bnd, g, s, t0, rst0, jw0, top, slot, k, kPrev, tM,
                                       rM, result);
      if (rho == rM) out(sp, result);
      ++sp;
      tM += pc;
    } else if (rst0 < G) {
      consume_period<A, G, true, false>(ring, bnd, g, s, t0, rst0, jw0, top, slot, k, kPrev, tM,
                                        rM, result);
    } else {
      consume_period<A, G, false, false>(ring, bnd, g, s, t0, rst0, jw0, top, slot, k, kPrev, tM,
                                         rM, result);
    }
    rst0 += kChunk;
    if (rst0 == g.Np) {
      rst0 = 0;
      if (++band0 == g.nb) band0 = 0;
    }
    jw0 += kChunk;
    if (jw0 >= g.Np) jw0 -= g.Np;
  }
}

}  // namespace sigpde
