// Fused fp16x2f linear-kernel Gram (fp16 hi/lo MFMA increments, float32 wavefront) on
// workgroups of NC consumer / producer wave pairs: producer p multiplies the increment tiles of
// its consumer's four pairs on MFMA into that consumer's LDS ring, consumer p runs the wavefront
// over them; all pairs of a workgroup share one stream of y paths, whose operands are fetched
// once per workgroup (the scheme of stojl/SigPDE sigpde/cuda_device_functions.py:3-5 and
// sigpde/cuda_gram_kernels.py:8-76, the increments of sigpde/torch/kernel.py:13-21 with
// sigpde/torch/static_kernels.py:7-11).
//
// Why split the roles: in one wave the producer's x / y operands and accumulators (about 100
// VGPRs at dim 64) hold occupancy at two waves per SIMD and the wave's in-order issue stalls the
// wavefront while its MFMAs drain; and every wave re-reading its y path for every band made the
// y operand loads L2-bandwidth bound. Split, a consumer wave carries only the wavefront's state,
// and a chunk's y operands are loaded once for NC pairs of waves (measured on the 1024 x 1024
// Gram of length 256, dim 64: 44.4 ms in one wave, 41.3 ms with NC = 1, 38.6 ms with NC = 4).
//
// Schedule. The wavefront is ring16f.h's (lane rho of a 16-lane group owns refined rows
// R rho .. R rho + R - 1 of a band of 16 R rows and takes virtual column v at step v + rho; four
// pairs per consumer, x paths 4 xq .. 4 xq + 3 against the stream of y paths; float32 cells in
// the split form of cell_f; the value above by v_add_f32 with a DPP row_shr:1 operand plus z, the
// boundary value on a group's first lane and 0 elsewhere). Chunk k (16 virtual columns, all 64
// ring rows) goes to ring slots (16 k + c + (r & 15)) mod 48 of ring row r; consumer period p
// (steps 16 p .. 16 p + 15) reads chunks p - 1 and p, each consumer group's inputs loaded one
// group ahead (so chunk k is first loaded in group 3 of period k - 1 and last loaded in group 2
// of period k + 1). Workgroup barriers: beta_-1 before the first period, beta_p between groups 2
// and 3 of period p (its LDS wait completes the loads issued before it). A producer writes chunk
// k between beta_(k-2) and beta_(k-1): after the last loads of chunk k - 3, whose slots it takes,
// and before the first loads of chunk k. The y operands of chunk k travel one window ahead of
// that: loaded (each producer its share) after beta_(k-5), staged in LDS after beta_(k-4), read
// back and multiplied after beta_(k-3). A band's x operands are loaded right after the band's
// last chunk is multiplied (one window ahead of their use).
//
// Ring layout: row-major, ring row r at physical row g16f_prow(r), kG16FRowStride floats per row,
// so a consumer group's four slots are two 8-byte LDS reads at immediate offsets; the row
// permutation puts the producer's chunk writes two lanes to a bank (three in one of the three
// slot phases, where 48-slot wrap-around shifts some lanes by 16 banks) and keeps the consumer's
// 8-byte reads conflict-free.
//
// Results are bit for bit those of ring16f.h's single-wave solve with the same producer
// (MfmaProducer<KB, 4, 1, kPrecFp16x2F>), for every NC: the same tiles (same MFMA order into one
// accumulator) and the same cell arithmetic in the same order.
#pragma once

#include "mfma_producer.h"

namespace sigpde {

constexpr int kG16FSlots = 48;      // ring slots: skew 16 + two chunks
constexpr int kG16FRowStride = 50;  // floats per ring row (48 slots + 2 pad)

// Physical ring row of ring row L (16 g + 4 h + e): 16 g + 9 (K[e] + H[h]) mod 16 with
// K = {0, 1, 8, 9}, H = {0, 6, 12, 2} (a permutation of each group's 16 rows).
__host__ __device__ inline int g16f_prow(int L) {
  const int e = L & 3, h = (L >> 2) & 3;
  const int K = (e & 1) | ((e & 2) << 2);  // 0, 1, 8, 9
  const int H = (6 * h) & 15;              // 0, 6, 12, 2
  return (L & 48) | ((9 * (K + H)) & 15);
}

// Floats of a consumer's sink: the line writes of lanes other than a group's last land there, at
// lane + 0 .. 15, so the writes need no branch on the lane.
constexpr int kG16FSink = 80;

// LDS of one consumer: four boundary lines (r16_line_stride<MixedF>), the zero block, the ring,
// the sink.
__host__ __device__ inline size_t g16f_lds_bytes(int N) {
  return r16_line_bytes<Arith::MixedF>(N) +
         ((size_t)kWave * kG16FRowStride + kG16FSink) * sizeof(float);
}

// ------------------------------------------------------------------ producer waves
// Producer p of a workgroup multiplies the tiles of its consumer's four pairs (x paths xa[0..3]
// against the workgroup's y stream). The y operands of a chunk (2 KB loads of 1 KB: hi / lo part
// of K block kb, each lane's 16 bytes in MFMA operand order) are fetched once per workgroup: load
// j goes to producer j mod NC, lands in its registers (ld), and is written to an LDS staging slot
// the next window; every producer reads the whole chunk back from the slot before multiplying it.
template <int KB, int NC>
struct G16FProducer {
  static constexpr int G = 16;
  static constexpr int NLD = 2 * KB;              // 1 KB loads per chunk
  static constexpr int NL = (NLD + NC - 1) / NC;  // loads per producer (upper bound)
  const _Float16 *xh, *xl, *yh, *yl;  // split operands: rows of dpad halves per increment
  int dpad, m, n;
  int p;                               // producer index in the workgroup
  int64_t xa[4];                       // x path of each pair (-1: idle, served path 0)
  int64_t ybase;                       // y path of stream position 0
  int64_t ny;                          // valid y paths (later positions clamp to ny - 1)
  half8 Ah[4][KB], Al[4][KB];          // the band's x operands
  half8 Bh[KB], Bl[KB];                // the y operands of the chunk being multiplied
  half8 ld[NL];                        // this producer's loads of a chunk, one window in flight
  float4v acc[4];                      // the last chunk multiplied, one 16 x 16 tile per pair
  int64_t yrow;                        // first element of the y path being fetched
  Cursor<G> cB, cC;                    // next chunk to fetch / to multiply
  uint32_t woff[4];                    // LDS address of element e's slot at phase 0 (pair 0)
  uint32_t woff2[4];                   // the same at phase 2 (wrapped past the ring's last slot)

  __device__ __forceinline__ uint32_t kslice() const { return 8 * ((threadIdx.x & 63) >> 4); }

  // ring: the consumer's ring in LDS (its LDS address goes into woff / woff2)
  __device__ __forceinline__ void init_writer(const float *ring) {
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, h = lane >> 4;
    const uint32_t rb = (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) float *)ring;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int s = c + 4 * h + e;  // slot at phase 0 (<= 30)
      woff[e] = rb + (uint32_t)((g16f_prow(4 * h + e) * kG16FRowStride + s) * sizeof(float));
      woff2[e] = woff[e] + 2 * kChunk * sizeof(float) -
                 (s >= kG16FSlots - 32 ? kG16FSlots * sizeof(float) : 0);
    }
  }
  __device__ __forceinline__ void fetchA(const PairGeom<G> &geo, int band) {
    const int c = threadIdx.x & 15;
    int ro = (band * G * geo.R + geo.R * c) >> geo.order;
    ro = ro > m - 1 ? m - 1 : ro;
    const uint32_t lo = (uint32_t)(ro * dpad) + kslice();
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t xrow = (xa[a] < 0 ? 0 : xa[a]) * m * (int64_t)dpad;
#pragma unroll
      for (int v = 0; v < KB; ++v) {
        Ah[a][v] = *reinterpret_cast<const half8 *>(xh + xrow + lo + v * kKBlock);
        Al[a][v] = *reinterpret_cast<const half8 *>(xl + xrow + lo + v * kKBlock);
      }
    }
  }
  // element offset of this lane's 16 bytes of chunk cB in a y operand row (yrow set at a pair's
  // first chunk)
  __device__ __forceinline__ uint32_t b_offset(const PairGeom<G> &geo) {
    if (cB.band == 0 && cB.jv0 == 0) {
      int64_t yp = ybase + cB.sp;
      yp = yp < ny ? yp : ny - 1;
      yrow = yp * n * dpad;
    }
    int co = (cB.jv0 + (threadIdx.x & 15) - 1) >> geo.order;
    co = co < 0 ? 0 : (co > n - 1 ? n - 1 : co);
    return (uint32_t)(co * dpad) + kslice();
  }
  // the whole of chunk cB into Bh / Bl (the first two chunks), then advance cB
  __device__ __forceinline__ void fetchB_all(const PairGeom<G> &geo) {
    const uint32_t lo = b_offset(geo);
#pragma unroll
    for (int v = 0; v < KB; ++v) {
      Bh[v] = *reinterpret_cast<const half8 *>(yh + yrow + lo + v * kKBlock);
      Bl[v] = *reinterpret_cast<const half8 *>(yl + yrow + lo + v * kKBlock);
    }
    cB.advance(geo);
  }
  // this producer's loads j = p, p + NC, ... of chunk cB into ld, then advance cB
  __device__ __forceinline__ void fetchB_share(const PairGeom<G> &geo) {
    const uint32_t lo = b_offset(geo);
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const int j = p + t * NC;
      if (j < NLD)
        ld[t] = *reinterpret_cast<const half8 *>((j & 1 ? yl : yh) + yrow + lo + (j >> 1) * kKBlock);
    }
    cB.advance(geo);
  }
  // ld -> staging slot: load j at slot + j KB, each lane's 16 bytes at 16 lane
  __device__ __forceinline__ void stage(char *slot) const {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const int j = p + t * NC;
      if (j < NLD) *reinterpret_cast<half8 *>(slot + j * 1024 + 16 * lane) = ld[t];
    }
  }
  __device__ __forceinline__ void unstage(const char *slot) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < KB; ++v) {
      Bh[v] = *reinterpret_cast<const half8 *>(slot + (2 * v) * 1024 + 16 * lane);
      Bl[v] = *reinterpret_cast<const half8 *>(slot + (2 * v + 1) * 1024 + 16 * lane);
    }
  }
  // The tiles of chunk cC into acc: hi*hi, then lo*hi + hi*lo, in one accumulator per pair
  // (MfmaProducer's kPrecFp16x2F order); then advance cC and, when the next chunk opens a band,
  // load that band's x operands.
  __device__ __forceinline__ void compute(const PairGeom<G> &geo, bool more) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      float4v c = {0, 0, 0, 0};
#pragma unroll
      for (int kb = 0; kb < KB; ++kb)
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[g][kb], Bh[kb], c, 0, 0, 0);
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[g][kb], Bh[kb], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[g][kb], Bl[kb], c, 0, 0, 0);
      }
      acc[g] = c;
    }
    cC.advance(geo);
    if (more && cC.jv0 == 0) fetchA(geo, cC.band);
  }
  // acc -> ring slots of chunk phase ph (chunk index mod 3): per element one address, the four
  // pairs' slots at immediate offsets from it
  __device__ __forceinline__ void store(int ph) const {
    typedef __attribute__((address_space(3))) float lds_float;
    const uint32_t po = (uint32_t)(ph * kChunk * sizeof(float));
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t a = ph == 2 ? woff2[e] : woff[e] + po;
#pragma unroll
      for (int g = 0; g < 4; ++g)
        *(lds_float *)(uintptr_t)(a + g * 16 * kG16FRowStride * sizeof(float)) = acc[g][e];
    }
  }
};

// LDS of the y operand staging: two slots of a chunk
template <int KB>
__host__ __device__ constexpr size_t g16f_stage_bytes() { return (size_t)2 * 2 * KB * 1024; }

// LDS of a workgroup of NC consumer / producer pairs: the staging, then per consumer its four
// boundary lines, zero block and ring.
template <int KB, int NC>
__host__ __device__ inline size_t g16f_wg_lds_bytes(int N) {
  return g16f_stage_bytes<KB>() + (size_t)NC * g16f_lds_bytes(N);
}

// A producer wave's loop. Chunk j is loaded (each producer its share) in window j - 4, staged in
// window j - 3, multiplied by every producer in window j - 2 and written to its consumer's ring in
// window j - 1 (window i: after beta_(i-1)); the first chunks are loaded whole by every producer.
// A window stages the loads of the previous window first and then issues its own, so a load has
// the rest of its window (the ring writes and the multiplications) to arrive.
// A staging slot holds chunk j from window j - 3 to window j - 2, and chunk j + 2 is staged in
// window j - 1, after the barrier that ends the reads of chunk j. nperiods + 1 barriers, as many
// as each consumer's. The ring starts zeroed (slots of virtual columns before the stream are read
// as c = 0, as in ring16f_solve).
template <int KB, int NC>
__device__ void g16f_produce(G16FProducer<KB, NC> &P, const PairGeom<16> &geo, float *ring,
                             char *stage) {
  constexpr size_t kSlot = g16f_stage_bytes<KB>() / 2;
  const int lane = threadIdx.x & 63;
  const int nchunks = geo.columns() / kChunk;
  const int nperiods = nchunks + 1;
  for (int i = lane; i < kWave * kG16FRowStride; i += 64) ring[i] = 0.0f;
  P.init_writer(ring);
  P.cB.start(geo, 0);
  P.cC.start(geo, 0);
  P.fetchA(geo, 0);
  P.fetchB_all(geo);                     // chunk 0
  P.compute(geo, nchunks > 1);
  P.store(0);
  if (nchunks > 1) {
    P.fetchB_all(geo);                   // chunk 1
    P.compute(geo, nchunks > 2);
  }
  if (nchunks > 2) {                     // chunk 2: staged in slot 0 before beta_-1
    P.fetchB_share(geo);
    P.stage(stage);
  }
  if (nchunks > 3) P.fetchB_share(geo);  // chunk 3: staged in window 0
  int ph = 1;  // phase (index mod 3) of the chunk written next
  int i = 0;
  // windows in which every step applies (the loop below, without its tests)
  for (; i + 4 < nchunks; ++i) {
    __syncthreads();  // beta_(i-1)
    const int k = i + 1;
    P.stage(stage + ((k + 2) & 1) * kSlot);
    P.fetchB_share(geo);
    P.store(ph);
    P.unstage(stage + ((k + 1) & 1) * kSlot);
    P.compute(geo, true);
    ph = ph == 2 ? 0 : ph + 1;
  }
  for (; i <= nperiods; ++i) {
    __syncthreads();  // beta_(i-1)
    const int k = i + 1;  // acc holds chunk k
    if (k + 2 < nchunks) P.stage(stage + ((k + 2) & 1) * kSlot);  // chunk i + 3
    if (k + 3 < nchunks) P.fetchB_share(geo);                       // chunk i + 4
    if (k < nchunks) P.store(ph);
    if (k + 1 < nchunks) {                                          // chunk i + 2
      P.unstage(stage + ((k + 1) & 1) * kSlot);
      P.compute(geo, k + 2 < nchunks);
    }
    ph = ph == 2 ? 0 : ph + 1;
  }
}

// ------------------------------------------------------------------ consumer wave
template <int R, int I, class F>
__device__ __forceinline__ void g16f_group(const char *rb, const R16FPeriod &pp, R16FState<R> &st,
                                           R16FIn &nxt, double &result, const char *rbNext) {
  const int rho = threadIdx.x & 15;
  R16FIn cur = nxt;
  if constexpr (I < 3)
    r16f_load(rb, 4 * (I + 1), pp.z + 4 * (I + 1), nxt);
  else
    r16f_load(rbNext, 0, pp.zNext, nxt);
  float vals[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int sidx = 4 * I + k;
    float a, b;
    coefs_f(cur.c[k], st.h, st.q, a, b);
    const float up = shift_add_f(st.k[R - 1], cur.z[k]);
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
  if constexpr (!F::kReset) {
    // every lane writes, without a branch: a group's last lane into its line, the others into
    // their sink slots (pp.w)
    float *w = pp.w + 4 * I;
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = F::kOnes ? 1.0f : vals[k];
  } else if (rho == 15) {
    // reset / capture periods: the line (ones after a pair's last band) and, in periods that
    // start a band, its mirror
    float *w = pp.w + 4 * I;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float v = pp.ones ? 1.0f : vals[k];
      w[k] = v;
      if (pp.wm) pp.wm[4 * I + k] = v;
    }
  }
}

template <int R, class F>
__device__ __forceinline__ void g16f_period(const char *rb, const R16FPeriod &pp, R16FState<R> &st,
                                            R16FIn &nxt, double &result, const char *rbNext) {
  g16f_group<R, 0, F>(rb, pp, st, nxt, result, rbNext);
  g16f_group<R, 1, F>(rb, pp, st, nxt, result, rbNext);
  g16f_group<R, 2, F>(rb, pp, st, nxt, result, rbNext);
  __syncthreads();  // beta_p: the next period's first chunk is in the ring after it
  g16f_group<R, 3, F>(rb, pp, st, nxt, result, rbNext);
}

// The consumer wave: ring16f_solve's stream without a producer; barriers beta_-1 before the
// first period and beta_p between groups 2 and 3 of period p. pair_scale(sp): the scale m of stream pair sp for this lane's group (c = m p,
// read one pair ahead of its first reset); out(sp, value) as in ring16f_solve.
template <int R, class Scale, class Out>
__device__ void g16f_consume(const PairGeom<16> &g, float *bnd, float *ring, Scale &&pair_scale,
                             Out &&out) {
  const int lane = threadIdx.x & 63;
  const int rho = lane & 15;
  const int Np = g.Np;
  const int bs = r16_line_stride<Arith::MixedF>(Np);
  float *line = bnd + (lane >> 4) * bs;
  const float *zeros = bnd + 4 * bs;
  const char *rb0 = reinterpret_cast<const char *>(ring + g16f_prow(lane) * kG16FRowStride);
  float *sink = ring + kWave * kG16FRowStride;  // kG16FSink floats after the ring

  const int nchunks = g.columns() / kChunk;
  const int nperiods = nchunks + 1;  // the last lane finishes at step columns + 14
  const int pc = g.pair_columns();
  const int bM = (g.M - 1) / (16 * R);
  const int rM = ((g.M - 1) - bM * 16 * R) / R;
  int tM = bM * Np + g.N + rM;  // result step of stream pair sp
  int sp = 0;

  R16FState<R> st;
#pragma unroll
  for (int r = 0; r < R; ++r) st.k[r] = 1.0f;
  st.kPrev = 1.0f;
  st.h = st.q = st.hN = st.qN = 0.0f;
  const float *zbase = rho == 0 ? line + 16 : zeros;
  const int zstep = rho == 0 ? 1 : 0;
  double mNext = pair_scale(0);  // scale of the next pair to start
  int npair = 0;

  R16FIn nxt;
  __syncthreads();  // beta_-1: chunk 0 is in the ring
  r16f_load(rb0, 0, zbase, nxt);
  int rst0 = 0;   // t0 mod Np
  int band = 0;   // band of column t0 within its pair
  int ph = 0;     // period index mod 3 (ring slots 16 ph .. 16 ph + 15)
  for (int per = 0; per < nperiods; ++per) {
    const char *rb = rb0 + ph * kChunk * sizeof(float);
    const int phn = ph == 2 ? 0 : ph + 1;
    const char *rbNext = rb0 + phn * kChunk * sizeof(float);
    const bool reset = rst0 == 0;
    const int rn = rst0 + kChunk == Np ? 0 : rst0 + kChunk;
    R16FPeriod pp;
    pp.z = zbase + zstep * rst0;
    pp.zNext = zbase + zstep * rn;
    pp.w = rho == 15 ? line + rst0 + 1 : sink + lane;  // reset periods: only rho 15 writes
    pp.wm = reset ? line + Np + 1 : nullptr;
    pp.ones = reset ? band == 0 : band == g.nb - 1;
    const bool cap = sp < g.stream && (tM >> 4) == per;
    pp.sCap = cap ? tM & 15 : -1;
    pp.rsi = reset ? rho : -1;
    if (reset && band == 0) {
      const double m = mNext;
      st.hN = (float)(0.5 * m);
      st.qN = (float)(m * m * (1.0 / 12));
      mNext = pair_scale(++npair);
    }
    double result = 0.0;
    const int fl = (reset || cap) ? 2 : (pp.ones ? 1 : 0);
    if (fl == 2)
      g16f_period<R, R16Special>(rb, pp, st, nxt, result, rbNext);
    else if (fl == 1)
      g16f_period<R, R16Ones>(rb, pp, st, nxt, result, rbNext);
    else
      g16f_period<R, R16Plain>(rb, pp, st, nxt, result, rbNext);
    if (cap) {
      if (rho == rM) out(sp, result);
      ++sp;
      tM += pc;
    }
    rst0 = rn;
    if (rn == 0 && ++band == g.nb) band = 0;
    ph = phn;
  }
}

// One work unit on a workgroup of 2 NC waves: waves 0 .. NC-1 are consumers (consumer p solves the
// four pairs of producer p), waves NC .. 2NC-1 the producers; all share one y stream. lds:
// g16f_wg_lds_bytes<KB, NC>(N) of the workgroup's LDS. Every wave must call it with the same
// geometry.
template <int R, int KB, int NC, class Scale, class Out>
__device__ void g16f_solve(G16FProducer<KB, NC> &P, const PairGeom<16> &g, void *lds,
                           Scale &&pair_scale, Out &&out) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int pi = wave < NC ? wave : wave - NC;
  const int bs = r16_line_stride<Arith::MixedF>(g.Np);
  char *stage = reinterpret_cast<char *>(lds);
  float *bnd = reinterpret_cast<float *>(stage + g16f_stage_bytes<KB>() +
                                         (size_t)pi * g16f_lds_bytes(g.N));
  float *ring = bnd + 4 * bs + kR16FZeros;
  if (wave < NC) {
    for (int i = lane; i < 4 * bs; i += 64) bnd[i] = 1.0f;
    if (lane < kR16FZeros) bnd[4 * bs + lane] = 0.0f;
    g16f_consume<R>(g, bnd, ring, pair_scale, out);
  } else {
    P.p = pi;
    g16f_produce<KB, NC>(P, g, ring, stage);
  }
}

}  // namespace sigpde
