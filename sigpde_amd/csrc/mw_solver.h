// Goursat PDE wavefront of ONE pair on W waves of a workgroup (latency path for few, long pairs).
//
// Same scheme and cell arithmetic as ring_solver.h / ring16.h (stojl/SigPDE
// sigpde/cuda_device_functions.py:3-5, sigpde/cuda_pairwise_kernels.py:8-39): lane rho of a wave
// owns refined rows 64 R b + R rho + 1 .. + R of band b (R = 1, or 2 from dyadic order 1: rows
// that share one row of increments) and handles column v = tau - rho + 1 at band step tau; the
// value above arrives from lane rho - 1 by DPP (wave_shr:1), and the same coefs / cell functions
// round every cell as the other solvers do, so a pair's result is the same bit for bit.
//
// What differs is that the bands of one pair run concurrently on the W waves, pipelined:
//   * Band b runs on wave b mod W during periods (16 steps each) [b D, b D + P), P = ceil((N+63)/16)
//     periods per band. Every wave steps through the same global period counter and the workgroup
//     synchronises once per period, so the schedule is static: no flags, no spinning.
//   * Band b's last row reaches band b+1's first lane through an LDS window owned by the consumer
//     wave: lane 63 writes column v at index (v + 63) mod win, lane 0 reads it D periods later
//     (its 16 columns of a period are then one aligned block of 16 slots; the first band reads a
//     window of ones, the boundary k[0][j] = 1, filled at the start of every solve). D >= 5
//     covers the 63-step skew of a band (the value of column v is written in period
//     b D + floor((v + 62) / 16) and read in period (b+1) D + floor((v - 1) / 16)); win = 16 (D + 2)
//     entries keep a value until it is read; (W - 1) D >= P + 1 (D's definition) makes a window
//     free again (its reader done; band 0's window of ones included) before the band W later
//     starts writing into it.
//   * Increments are read straight from global memory (cached: a path's increments are reused by
//     every root-find iteration), a period ahead of their use: a run of 16 columns of the lane's
//     row per period.
// Sequential steps per pair: ((nb - 1) D + P) 16 instead of nb P 16 on one wave
// (nb = ceil(M / (64 R))).
//
// Spread over G workgroups (XW; for fewer pairs than the GPU holds workgroups): the WT = G W
// waves of a pair's workgroups share its bands as one workgroup's W waves do (band b on wave
// b mod WT of workgroup (b mod WT) / W; D from WT). Inside a workgroup the handoff stays in LDS;
// between workgroups (wave W - 1 of workgroup g feeds wave 0 of workgroup g + 1 mod G) it goes
// through a window in global memory: link l (into workgroup l) has two buffers of wing =
// 16 (D + 2 + kXwgSlack) entries, used by alternate rounds of WT bands. The workgroups run
// their periods on their own, each publishing the count of periods it has completed (release,
// agent scope) at the top of every period, and wait (acquire) only where the schedule needs a
// neighbour's progress:
//   * reading (wave 0, band b >= 1, period p): columns up to 16 q + 16 were written by band b - 1
//     by its period q + 4, so the feeding workgroup must have completed p - D + 5 periods;
//   * writing (wave W - 1, period p): a slot's earlier value was read at least 5 + kXwgSlack
//     periods before (the same band's column v - wing, or the round two bands-of-WT before), so
//     the fed workgroup must have completed p - 4 - kXwgSlack periods.
// Both needs are below the waiting workgroup's own period, so the workgroup furthest behind never
// waits: with all of a pair's workgroups resident, every wait ends. Every wait is still bounded
// (wall clock); one that gives up sets the launch's failure flag, after which no wait waits, the
// kernel runs to its end, and the host reports the launch as failed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "ring16.h"
#include "ring_solver.h"

namespace sigpde {

struct MwGeom {
  int m, n, order;  // original increments (rows / columns) and dyadic order
  int M, N;         // refined rows / columns
  int R;            // refined rows per lane (a power of two dividing 2^order: they share a row
                    // of increments)
  int nb;           // bands of 64 R rows
  int P;            // periods per band
  int D;            // periods between band starts
  int win;          // window entries
  int T;            // periods of the whole pair
  // W: the waves that share the pair's bands (all workgroups of a spread pair); Dmin: the least
  // band offset (5 covers a band's skew; a pair spread over workgroups takes more as slack)
  __host__ __device__ MwGeom(int m_, int n_, int o, int W, int R_ = 1, int Dmin = 5)
      : m(m_), n(n_), order(o), R(R_) {
    M = m << o;
    N = n << o;
    nb = (M + 64 * R - 1) / (64 * R);
    P = (N + 63 + 15) / 16;
    const int need = (P + 1 + (W - 2)) / (W - 1);  // ceil((P + 1) / (W - 1))
    D = need > Dmin ? need : Dmin;
    win = 16 * (D + 2);
    T = (nb - 1) * D + P;
  }
};

// The 16 consecutive values p[0..15] (p aligned to its element), as 16-byte vector loads.
template <typename RT>
__device__ __forceinline__ void mw_load_run16(const RT *p, RT (&dst)[16]) {
  constexpr int E = 16 / sizeof(RT);
  typedef RT vec_t __attribute__((ext_vector_type(E), aligned(sizeof(RT))));
#pragma unroll
  for (int k = 0; k < 16 / E; ++k) {
    const vec_t v = *reinterpret_cast<const vec_t *>(p + E * k);
#pragma unroll
    for (int e = 0; e < E; ++e) dst[E * k + e] = v[e];
  }
}

// dst[st] = row[(c0 + st) >> O] for st = 0..15 (O >= 1, c0 >= 0, row[(c0 + 15) >> O] in range):
// the distinct values are loaded once each and repeated. Every index is a constant of a fold
// expression (no loops), so the values stay in registers: a loop, before it is unrolled, would
// let the compiler turn the per-lane choice between two values into an indexed scratch access.
template <int O, typename RT, int... K, int... ST>
__device__ __forceinline__ void mw_load_rep16(const RT *row, int c0, RT (&dst)[16],
                                              std::integer_sequence<int, K...>,
                                              std::integer_sequence<int, ST...>) {
  constexpr int kMask = (1 << O) - 1;
  constexpr int kCnt = sizeof...(K);
  static_assert(kCnt == ((kMask + 15) >> O) + 1, "distinct values of 16 columns");
  static_assert((15 >> O) + 1 < kCnt, "index of the last step's value");
  const int j0 = c0 >> O, jl = (c0 + 15) >> O, ph = c0 & kMask;
  const RT v[kCnt] = {row[j0 + K < jl ? j0 + K : jl]...};
  // (ph + st) >> O = (st >> O) + 1 exactly when ph + (st & kMask) passes kMask
  ((dst[ST] = ph + (ST & kMask) > kMask ? v[(ST >> O) + 1] : v[ST >> O]), ...);
}

template <int O, typename RT>
__device__ __forceinline__ void mw_load_rep16(const RT *row, int c0, RT (&dst)[16]) {
  mw_load_rep16<O>(row, c0, dst, std::make_integer_sequence<int, (((1 << O) + 14) >> O) + 1>(),
                   std::make_integer_sequence<int, 16>());
}

// Periods of slack in a cross-workgroup window (the writer may run this far ahead of the reader).
constexpr int kXwgSlack = 8;

// One workgroup's view of a pair spread over G >= 2 workgroups. Every thread holds a copy;
// seenIn / seenOut / dead are the same on every lane of a wave.
struct Xwg {
  void *win;        // G links x 2 buffers x wing entries of the window type
  int *done;        // G counters: periods each workgroup of the pair has completed
  int *err;         // the launch's failure flag (nonzero: a wait gave up)
  long long limit;  // wall-clock ticks one wait may take
  int G, gi;        // workgroups of the pair, this one's index
  int wing;         // entries of a global window buffer
  int hook;         // 1: workgroup 0 publishes no progress (failure-path test)
  int base;         // periods of the pair's earlier solves (each solve runs T on every workgroup)
  int seenIn, seenOut;  // the largest counts this wave has synchronised with (acquire)
  bool dead;
};

__host__ __device__ inline int xwg_window_entries(int D) { return 16 * (D + 2 + kXwgSlack); }

// Waits until *flag >= need (seen: the largest value this wave has read; it needs no read while
// it covers need). Acquire, agent scope. Gives up after x.limit wall-clock ticks, or at once when
// the launch has failed, setting the failure flag; returns false then.
__device__ inline bool xwg_wait(const int *flag, int need, int &seen, Xwg &x) {
  if (seen >= need) return true;
  if (x.dead) return false;
  const long long t0 = (long long)wall_clock64();
  for (;;) {
    const int v = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (v >= need) {
      seen = v;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      return true;
    }
    const int failed = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(x.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (failed || (long long)wall_clock64() - t0 > x.limit) {
      __hip_atomic_store(x.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      x.dead = true;
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Accesses through pointers of a stated address space (3: LDS, 1: global), so that code choosing
// between an LDS and a global window keeps LDS and global instructions instead of flat ones.
template <typename V>
__device__ __forceinline__ V as_load(const void *p, std::integral_constant<int, 3>) {
  return *(const __attribute__((address_space(3))) V *)(p);
}
template <typename V>
__device__ __forceinline__ V as_load(const void *p, std::integral_constant<int, 1>) {
  return *(const __attribute__((address_space(1))) V *)(p);
}
template <typename V>
__device__ __forceinline__ void as_store(void *p, V v, std::integral_constant<int, 3>) {
  *(__attribute__((address_space(3))) V *)(p) = v;
}
template <typename V>
__device__ __forceinline__ void as_store(void *p, V v, std::integral_constant<int, 1>) {
  *(__attribute__((address_space(1))) V *)(p) = v;
}

// The workgroup (of a pair spread over G) whose waves hold the last band, and so K[M][N].
__host__ __device__ inline int mw_corner_group(const MwGeom &g, int W, int G) {
  return ((g.nb - 1) % (W * G)) / W;
}

// LDS of one workgroup: W windows of the band-boundary type, then one double (the result).
template <Arith A>
__host__ __device__ inline size_t mw_lds_bytes(const MwGeom &g, int W) {
  using BT = typename R16Bnd<A>::T;
  const size_t w = (size_t)W * g.win * sizeof(BT);
  return (w + 7) / 8 * 8 + sizeof(double);
}

// Solves one pair: increments inc (m x n, row stride ld), pair scale `scale` applied to each
// increment in RT when `scaled` (as the other producers do), Scaled* modes' per-call factor s.
// R refined rows per lane (g.R == R): lane rho of band b holds rows 64 R b + R rho + 1 .. + R,
// which share one row of increments, so the coefficients and the shift from the lane above serve
// R cells a step. Every thread of the workgroup (64 W threads) must call it; all get K[M][N].
//
// XW: the pair is spread over xs->G workgroups (g built with W G waves); the result is K[M][N]
// only in workgroup mw_corner_group(g, W, G).
template <Arith A, int W, int R, bool XW = false>
__device__ double mw_solve(const typename ArithTraits<A>::RT *__restrict__ inc, int64_t ld,
                           typename ArithTraits<A>::RT scale, bool scaled, const MwGeom &g,
                           double s, void *lds, Xwg *xs = nullptr) {
  using RT = typename ArithTraits<A>::RT;
  using BT = typename R16Bnd<A>::T;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  BT *wins = reinterpret_cast<BT *>(lds);
  double *res = reinterpret_cast<double *>(static_cast<char *>(lds) +
                                           ((size_t)W * g.win * sizeof(BT) + 7) / 8 * 8);
  BT *inWin = wins + wave * g.win;                    // band b-1 -> this wave's band b
  BT *outWin = wins + ((wave + 1) % W) * g.win;       // this wave's band b -> band b+1
  int inSz = g.win, outSz = g.win;
  const BT *gIn = nullptr;  // XW: wave 0's window in global memory (bands >= 1)
  BT *gOut = nullptr;       // XW: wave W - 1's window in global memory
  int WT = W, gw = wave;  // waves sharing the pair's bands, this wave's index among them
  if constexpr (XW) {
    WT = W * xs->G;
    gw = xs->gi * W + wave;
  }
  const int jo_max = g.n - 1;

  double k[R], kPrev = 1.0, result = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) k[r] = 1.0;
  // k of lane rho - 1 (lane 0: `top`, the band above). In the float32 modes every cell value is
  // a float (cell rounds it), so the float is shifted: one DPP move instead of two, no conversion
  // of the top value, the same value.
  auto shift_up = [](double kk, BT top) -> double {
    if constexpr (ArithTraits<A>::kRound) {
      static_assert(sizeof(BT) == 4, "float32 modes keep float windows");
      return shift_down_f<64>(kk, top);
    } else {
      return shift_down<64>(kk, (double)top);
    }
  };
  // one step of the lane's R rows at its current column: the row above the first is `up`
  // (lane rho - 1's last row), each further row's the one computed just before it
  auto step_rows = [&](double up, double cA, double cB) {
    double above = up, aboveOld = kPrev;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double old = k[r];
      k[r] = cell<A>(above, old, aboveOld, cA, cB);
      aboveOld = old;
      above = k[r];
    }
  };
  RT cc[16], cn[16];
  // increments of period q of band b, for this lane: columns v = 16 q + st - lane + 1. When all
  // 16 lie inside 1 .. N (16 q >= 63, 16 q + 16 <= N) they are one run of the lane's row: read as
  // vector loads at order 0 and as the few distinct values at higher orders, so that a wave's
  // load instructions touch each cache line a few times rather than once per column.
  auto fetch = [&](RT (&dst)[16], int b, int q) {
    int io = ((64 * R * b + R * lane) >> g.order);
    io = io < g.m ? io : g.m - 1;
    const RT *row = inc + (int64_t)io * ld;
    const int c0 = 16 * q - lane;  // v - 1 of step 0
    if (16 * q >= 63 && 16 * q + 16 <= g.N) {
      switch (g.order) {
        case 0: mw_load_run16(row + c0, dst); break;
        case 1: mw_load_rep16<1>(row, c0, dst); break;
        case 2: mw_load_rep16<2>(row, c0, dst); break;
        case 3: mw_load_rep16<3>(row, c0, dst); break;
        default: {  // 16 columns span at most two original columns
          const int j0 = c0 >> g.order, jl = (c0 + 15) >> g.order;
          const RT v0 = row[j0], v1 = row[jl];
#pragma unroll
          for (int st = 0; st < 16; ++st) dst[st] = ((c0 + st) >> g.order) == j0 ? v0 : v1;
        }
      }
    } else {
#pragma unroll
      for (int st = 0; st < 16; ++st) {
        int jo = (c0 + st) >> g.order;
        jo = jo < 0 ? 0 : (jo > jo_max ? jo_max : jo);
        dst[st] = row[jo];
      }
    }
    if (scaled) {
#pragma unroll
      for (int st = 0; st < 16; ++st) dst[st] = (RT)(dst[st] * scale);
    }
  };

  // band 0 (wave 0) reads the boundary row of ones from its window; the band that writes this
  // window next (band W - 1, from period (W - 1) D >= P) starts after band 0 has ended
  for (int x = threadIdx.x; x < g.win; x += 64 * W) wins[x] = (BT)1;
  __syncthreads();

  for (int p = 0; p < g.T; ++p) {
    if constexpr (XW) {
      // the periods before p are complete: wave W - 1's window writes of them (its lane 63
      // publishes, after its own stores) and wave 0's window reads (done before the barrier)
      if (p > 0 && wave == W - 1 && lane == 63 && !(xs->hook == 1 && xs->gi == 0))
        __hip_atomic_store(xs->done + xs->gi, xs->base + p, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // this wave's band in period p (at most one: WT D >= P + 1)
    int b = -1, q = 0;
    const int rel = p - gw * g.D;
    if (rel >= 0) {
      const int bb = gw + WT * (rel / (WT * g.D));
      const int qq = p - bb * g.D;
      if (bb < g.nb && qq < g.P) {
        b = bb;
        q = qq;
      }
    }
    if (b >= 0) {
      if constexpr (XW) {
        const int G = xs->G;
        if (wave == 0) {
          // band 0 reads the boundary row, ones, from wave 0's LDS window (which no band
          // writes when the pair is spread); the others read their link
          gIn = nullptr;
          inSz = g.win;
          if (b > 0) {
            gIn = static_cast<const BT *>(xs->win) +
                  ((size_t)xs->gi * 2 + ((b / WT) & 1)) * xs->wing;
            inSz = xs->wing;
            xwg_wait(xs->done + (xs->gi + G - 1) % G, xs->base + p - g.D + 5, xs->seenIn, *xs);
          }
        }
        if (wave == W - 1) {
          gOut = static_cast<BT *>(xs->win) +
                 ((size_t)((xs->gi + 1) % G) * 2 + (((b + 1) / WT) & 1)) * xs->wing;
          outSz = xs->wing;
          if (b + 1 < g.nb)
            xwg_wait(xs->done + (xs->gi + 1) % G, xs->base + p - 4 - kXwgSlack, xs->seenOut,
                     *xs);
        }
      }
      if (q == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) k[r] = 1.0;  // k[i][0]
        kPrev = 1.0;                              // k[i-1][0]
        fetch(cc, b, 0);
      } else {
#pragma unroll
        for (int st = 0; st < 16; ++st) cc[st] = cn[st];
      }
      if (q + 1 < g.P) fetch(cn, b, q + 1);
      const int i = 64 * R * b + R * lane + 1;  // this lane's first row (1-based)
      // lane 63 writes its columns 16 q + st - 62 at window slots wb + st + 1 (win is a multiple
      // of 16: only the last may wrap)
      const int wb = (16 * q) % outSz;
      const int w15 = wb + 16 < outSz ? wb + 16 : 0;
      // lane 0's row above the band for the period's columns 16 q + 1 .. 16 q + 16, all written
      // in earlier periods (D >= 5), at slots r0 .. r0 + 15: read (broadcast to every lane) four
      // at a time, a group of four steps ahead of their use. Slots of columns past N hold stale
      // values; they reach only lane 0's cells past N, which are not computed.
      const int r0 = (16 * q + 64) % inSz;
      auto tops4 = [&](BT (&t)[4], int grp) {
        typedef BT vec4 __attribute__((ext_vector_type(4), aligned(16)));
        vec4 v;
        if constexpr (XW) {
          if (gIn)
            v = as_load<vec4>(gIn + r0 + 4 * grp, std::integral_constant<int, 1>());
          else
            v = as_load<vec4>(inWin + r0 + 4 * grp, std::integral_constant<int, 3>());
        } else {
          v = *reinterpret_cast<const vec4 *>(inWin + r0 + 4 * grp);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) t[e] = v[e];
      };
      auto put = [&](int slot, BT val) {
        if constexpr (XW) {
          if (gOut)
            as_store<BT>(gOut + slot, val, std::integral_constant<int, 1>());
          else
            as_store<BT>(outWin + slot, val, std::integral_constant<int, 3>());
        } else {
          outWin[slot] = val;
        }
      };
      BT tcur[4], tnxt[4];
      tops4(tcur, 0);
      const bool hand = lane == 63 && b + 1 < g.nb;  // this lane feeds the next band
      // Steady periods: every lane's 16 columns lie inside 1 .. N - 1 and every row inside the
      // grid, so no step needs a range test (and none holds the corner, column N).
      const bool steady = 16 * q >= 63 && 16 * q + 16 < g.N && 64 * R * (b + 1) <= g.M;
      if (steady) {
#pragma unroll
        for (int st = 0; st < 16; ++st) {
          if (st % 4 == 0 && st < 12) tops4(tnxt, st / 4 + 1);
          const double up = shift_up(k[R - 1], tcur[st % 4]);
          if (st % 4 == 3 && st < 15) {
#pragma unroll
            for (int e = 0; e < 4; ++e) tcur[e] = tnxt[e];
          }
          double cA, cB;
          coefs<A>(cc[st], s, 0.0, 0.0, cA, cB);
          step_rows(up, cA, cB);
          if (hand) put(st < 15 ? wb + st + 1 : w15, (BT)k[R - 1]);
          kPrev = up;
        }
      } else {
#pragma unroll
        for (int st = 0; st < 16; ++st) {
          if (st % 4 == 0 && st < 12) tops4(tnxt, st / 4 + 1);
          const int v = 16 * q + st - lane + 1;  // this lane's column (1-based)
          const double up = shift_up(k[R - 1], tcur[st % 4]);
          if (st % 4 == 3 && st < 15) {
#pragma unroll
            for (int e = 0; e < 4; ++e) tcur[e] = tnxt[e];
          }
          // M is a multiple of R: a lane's rows are all inside the grid or all past it
          if (v >= 1 && v <= g.N && i + R - 1 <= g.M) {
            double cA, cB;
            coefs<A>(cc[st], s, 0.0, 0.0, cA, cB);
            step_rows(up, cA, cB);
            if (hand) put(st < 15 ? wb + st + 1 : w15, (BT)k[R - 1]);
            if (v == g.N) {
#pragma unroll
              for (int r = 0; r < R; ++r)
                if (i + r == g.M) result = k[r];
            }
          }
          kPrev = up;
        }
      }
    }
    __syncthreads();
  }
  if constexpr (XW) {
    if (wave == W - 1 && lane == 63 && !(xs->hook == 1 && xs->gi == 0))
      __hip_atomic_store(xs->done + xs->gi, xs->base + g.T, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    xs->base += g.T;
  }
  // the corner lane publishes K[M][N]
  if ((g.M - 1 - 64 * R * (g.nb - 1)) / R == lane && gw == (g.nb - 1) % WT) *res = result;
  __syncthreads();
  const double r = *res;
  __syncthreads();  // the next call may overwrite the slot
  return r;
}

}  // namespace sigpde
