// Float32 Goursat wavefront (Arith::MixedF) for lane groups of 16, four pairs per wave: the
// ring16.h schedule (lane rho of a group owns refined rows R rho .. R rho + R - 1 of a band and
// handles virtual column v at step v + rho; bands and pairs follow each other as one stream of
// virtual columns, Np per band; stojl/SigPDE sigpde/cuda_pairwise_kernels.py:8-39 and
// sigpde/cuda_device_functions.py:3-5 for the scheme), with the LDS layout and the cross-lane
// shift arranged for the float cell so that a step costs the cell's VALU work and little else:
//   * Row-major ring: ring row r (consumer lane r) keeps slots 0..31 at floats r * kR16FRowStride
//     + slot; increment (row r, column v) is in slot (v + (r & 15)) mod 32, so at step t every
//     lane reads slot t mod 32 of its own row and a consumer group's four columns are two 8-byte
//     LDS reads at immediate offsets (no address arithmetic in the loop).
//   * The value above, k[row - 1] of the previous step, reaches lane rho from lane rho - 1 of its
//     16-lane row by DPP row_shr:1 with out-of-row sources read as zero, and is added to z in the
//     same instruction (v_add_f32 with a DPP operand): z is the band-boundary value on a group's
//     first lane and 0 on the others, which read a block of zeros instead of the line (their
//     line address is selected once per period). The DPP source needs no copy of the old value.
//   * Boundary lines shifted by one element: the group's last lane writes its value of step T at
//     index (T mod Np) + 1, the first lane reads at step T index ((T + 15) mod Np) + 1, written at
//     step T + 15 - Np; a group's four reads are one 16-byte aligned read. Indices Np+1..Np+16
//     mirror 1..16 (written in the periods that start a band), so a period's reads never wrap.
// Results equal ring16.h's MixedF path bit for bit (same cell arithmetic in the same order; the
// value above is the same float, x + 0 = x).
#pragma once

#include "ring16.h"

namespace sigpde {

// Lane rho receives lane rho - 1 of its 16-lane row (0 on the row's first lane), plus z.
__device__ __forceinline__ float shift_add_f(float v, float z) {
  const float up =
      __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
  return up + z;
}

struct R16FIn {
  float c[4];   // ring: the lane's increments of the group's four steps
  float z[4];   // boundary values (a group's first lane) or zeros
};

template <int R>
struct R16FState {
  float k[R];    // the lane's rows at its current column
  float kPrev;   // the row above the lane's first row, at the previous column
  float h, q;    // coefficients of the lane's current pair
  float hN, qN;  // coefficients taken at the lane's next reset
};

struct R16FPeriod {
  const float *z;      // this period's boundary reads (line + rst0 + 16, or the zero block)
  const float *zNext;  // the next period's first group
  float *w;            // line + rst0 + 1 (writes of the group's last lane)
  float *wm;           // line + Np + 1 + rst0 (mirror writes; reset periods only)
  bool ones;
  int sCap;
  int rsi;
};

// rb: this lane's ring row (bytes); slot0: first slot of the group (a multiple of 4).
__device__ __forceinline__ void r16f_load(const char *rb, int slot0, const float *zp, R16FIn &in) {
  typedef float float2v __attribute__((ext_vector_type(2)));
  typedef float float4v_ __attribute__((ext_vector_type(4)));
  const float2v a = *reinterpret_cast<const float2v *>(rb + slot0 * sizeof(float));
  const float2v b = *reinterpret_cast<const float2v *>(rb + (slot0 + 2) * sizeof(float));
  in.c[0] = a.x;
  in.c[1] = a.y;
  in.c[2] = b.x;
  in.c[3] = b.y;
  const float4v_ z = *reinterpret_cast<const float4v_ *>(zp);
  in.z[0] = z.x;
  in.z[1] = z.y;
  in.z[2] = z.z;
  in.z[3] = z.w;
}

template <int R, int PAR, class F, int I>
__device__ __forceinline__ void r16f_group(const char *rb, const R16FPeriod &pp, R16FState<R> &st,
                                           R16FIn &nxt, double &result) {
  const int rho = threadIdx.x & 15;
  R16FIn cur = nxt;
  if constexpr (I < 3)
    r16f_load(rb, 16 * PAR + 4 * (I + 1), pp.z + 4 * (I + 1), nxt);
  else
    r16f_load(rb, 16 * (1 - PAR), pp.zNext, nxt);
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
  if (rho == 15) {
    float *w = pp.w + 4 * I;
    if constexpr (F::kOnes) {
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = 1.0f;
    } else if constexpr (F::kReset) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float v = pp.ones ? 1.0f : vals[k];
        w[k] = v;
        if (pp.wm) pp.wm[4 * I + k] = v;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = vals[k];
    }
  }
}

template <int R, int PAR, class F>
__device__ __forceinline__ void r16f_head(const char *rb, const R16FPeriod &pp, R16FState<R> &st,
                                          R16FIn &nxt, double &result) {
  r16f_group<R, PAR, F, 0>(rb, pp, st, nxt, result);
  r16f_group<R, PAR, F, 1>(rb, pp, st, nxt, result);
  r16f_group<R, PAR, F, 2>(rb, pp, st, nxt, result);
}

// Solves the wave's stream as ring16_solve<Arith::MixedF, R, GL> does (same producer interface,
// same out(sp, value) calls). lds: r16_lds_bytes<Arith::MixedF>(N) bytes (GL: the ring only,
// r16_ring_bytes<Arith::MixedF>(); gline then holds this wave's four lines and zero block,
// r16_line_bytes<Arith::MixedF>(N) bytes). The producer writes with Ring16WriterF.
template <int R, bool GL = false, class Producer, class Out>
__device__ void ring16f_solve(Producer &P, const PairGeom<16> &g, void *lds, Out &&out,
                              float *gline = nullptr) {
  const int lane = threadIdx.x & 63;
  const int rho = lane & 15;
  const int Np = g.Np;
  const int bs = r16_line_stride<Arith::MixedF>(Np);
  float *bnd;
  float *ring;
  if constexpr (GL) {
    bnd = gline;
    ring = reinterpret_cast<float *>(lds);
  } else {
    bnd = reinterpret_cast<float *>(lds);
    ring = bnd + 4 * bs + kR16FZeros;
  }
  float *line = bnd + (lane >> 4) * bs;
  const float *zeros = bnd + 4 * bs;
  const char *rb = reinterpret_cast<const char *>(ring + lane * kR16FRowStride);

  for (int i = lane; i < kWave * kR16FRowStride; i += 64) ring[i] = 0.0f;
  for (int i = lane; i < 4 * bs; i += 64) bnd[i] = 1.0f;
  if (lane < kR16FZeros) bnd[4 * bs + lane] = 0.0f;

  const int total = g.columns();
  const int nperiods = total / kChunk + 1;  // the last lane finishes at step total + 14
  const int pc = g.pair_columns();
  const int bM = (g.M - 1) / (16 * R);
  const int rM = ((g.M - 1) - bM * 16 * R) / R;
  int tM = bM * Np + g.N + rM;  // result step of stream pair sp
  int sp = 0;

  P.prologue(ring, g);

  R16FState<R> st;
#pragma unroll
  for (int r = 0; r < R; ++r) st.k[r] = 1.0f;
  st.kPrev = 1.0f;
  st.h = st.q = st.hN = st.qN = 0.0f;
  // the first lane of each group reads its line, the others the zero block
  const float *zbase = rho == 0 ? line + 16 : zeros;
  const int zstep = rho == 0 ? 1 : 0;
  R16FIn nxt;
  r16f_load(rb, 0, zbase, nxt);

  int rst0 = 0;   // t0 mod Np
  int band = 0;   // band of column t0 within its pair
  for (int per = 0; per < nperiods; ++per) {
    const bool reset = rst0 == 0;
    const int rn = rst0 + kChunk == Np ? 0 : rst0 + kChunk;
    R16FPeriod pp;
    pp.z = zbase + zstep * rst0;
    pp.zNext = zbase + zstep * rn;
    pp.w = line + rst0 + 1;
    pp.wm = reset ? line + Np + 1 : nullptr;
    pp.ones = reset ? band == 0 : band == g.nb - 1;
    const bool cap = sp < g.stream && (tM >> 4) == per;
    pp.sCap = cap ? tM & 15 : -1;
    pp.rsi = reset ? rho : -1;
    if (reset && band == 0) {
      const double m = P.lane_scale();
      st.hN = (float)(0.5 * m);
      st.qN = (float)(m * m * (1.0 / 12));
    }
    double result = 0.0;
    const bool odd = per & 1;
    const int fl = (reset || cap) ? 2 : (pp.ones ? 1 : 0);
    P.early(ring, g, per);
    if (fl == 2) {
      if (odd) r16f_head<R, 1, R16Special>(rb, pp, st, nxt, result);
      else r16f_head<R, 0, R16Special>(rb, pp, st, nxt, result);
    } else if (fl == 1) {
      if (odd) r16f_head<R, 1, R16Ones>(rb, pp, st, nxt, result);
      else r16f_head<R, 0, R16Ones>(rb, pp, st, nxt, result);
    } else {
      if (odd) r16f_head<R, 1, R16Plain>(rb, pp, st, nxt, result);
      else r16f_head<R, 0, R16Plain>(rb, pp, st, nxt, result);
    }
    // chunk per+1 (parity of per+1) goes to the ring after the reads of group 3 were issued
    if (odd) P.template period<0>(ring, g, per);
    else P.template period<1>(ring, g, per);
    if (fl == 2) {
      if (odd) r16f_group<R, 1, R16Special, 3>(rb, pp, st, nxt, result);
      else r16f_group<R, 0, R16Special, 3>(rb, pp, st, nxt, result);
    } else if (fl == 1) {
      if (odd) r16f_group<R, 1, R16Ones, 3>(rb, pp, st, nxt, result);
      else r16f_group<R, 0, R16Ones, 3>(rb, pp, st, nxt, result);
    } else {
      if (odd) r16f_group<R, 1, R16Plain, 3>(rb, pp, st, nxt, result);
      else r16f_group<R, 0, R16Plain, 3>(rb, pp, st, nxt, result);
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
