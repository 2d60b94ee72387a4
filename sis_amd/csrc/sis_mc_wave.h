// Limit-rule Monte Carlo of the performance mode: one wavefront resolves
// one connected component of the formed-pair graph, with the component's
// pairs and nucleotides held in registers (ballots and lane shuffles; no
// LDS round trips in the loop).  The algorithm per iteration is that of
// src/force_bp_limit_triplet.F90:259-330 restricted to the component:
//   pick a nucleotide with excess uniformly (ascending id order), list its
//   formed pairs in list order, shuffle them with the reference's swap
//   sweep, choose the pair to delete by the Metropolis sweep on energies,
//   delete it and update the two counts.
// Restricting to a component keeps the distribution: the uniform choice
// among all excess nucleotides, conditioned on the component, is uniform in
// it, and deletions never change counts outside the component.
#pragma once
#include "sis_device.h"

namespace sis {

constexpr int MCW_ITEMS = 2;   // pairs per lane   -> 128 pairs per component
constexpr int MCW_NODES = 4;   // nodes per lane   -> 256 nucleotides
// ints of wave scratch (LDS) for `items` pairs per lane
constexpr int mcw_lds(int items) { return 6 * 64 * items + 64; }
constexpr int MCW_LDS = mcw_lds(MCW_ITEMS);
// the large form: components of up to 256 pairs (the bucket capacity)
constexpr int MCB_ITEMS = 4, MCB_NODES = 8;
constexpr int MCB_LDS = mcw_lds(MCB_ITEMS);

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double philox_u(uint64_t seed, int r, int64_t step, int root, uint32_t draw) {
  U4 o = philox(U4{(uint32_t)step, (uint32_t)(step >> 32) ^ 0x4d43u, (uint32_t)root, draw},
                (uint32_t)seed, (uint32_t)(seed >> 32) ^ (0x5bd1e995u * (uint32_t)(r + 1)));
  return u01(o.x, o.y);
}

// Broadcasts from a lane chosen by a wave-uniform index go through
// v_readlane, and writes to one lane are a per-lane select: register
// traffic, no LDS round trip.
__device__ __forceinline__ int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
__device__ __forceinline__ int setlane(int v, int x, int lane) { return __lane_id() == lane ? x : v; }
__device__ __forceinline__ unsigned long long bcast64(unsigned long long v, int lane) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}
// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int L = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(v, d);
    if (L >= d) v += t;
  }
  return v;
}

// Set bits of m below this lane (prefix population count).
__device__ __forceinline__ int below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
}
// Orders LDS writes of this wave before its later LDS reads of other lanes' words.
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// The Monte Carlo loop of mc_component_wave over its pair slots (IT chunks
// of 64) and distinct nucleotides (ND chunks of 64), then the write-back of
// the deleted pairs.  Uniform arguments; all 64 lanes call it.
template <int IT, int ND>
__device__ __forceinline__ bool mcg_sweep(const Params &P, const Ptrs &D, int rk, int root, int n, int nn,
                                          double beta, uint64_t seed, int64_t step, int *sseq, const int (&gb)[IT],
                                          const int (&ia)[IT], const int (&ib)[IT], const double (&en)[IT],
                                          bool (&alive)[IT], const int (&nid)[ND], int (&ncnt)[ND]) {
  constexpr int NP = 64 * IT;
  const int L = lane_id();
  // ---- the Monte Carlo loop.  Draws come from a window of 64 counters,
  // lane i holding draw w0 + i and its logarithm (for the Metropolis test,
  // log u < dE beta, the same test as u < exp(dE beta)); the window moves
  // on when an iteration could run past it.
  uint32_t draw = 0, w0 = 0;
  double wu = philox_u(seed, rk, step, root, (uint32_t)L), wl = log(wu);
  // draws above one half: the swap target of a two-pair sequence is 1 there
  unsigned long long g2 = __ballot(wu > 0.5);
  const int nchunk_nodes = (nn + 63) / 64;
  const int nchunk_items = (n + 63) / 64;
  // live pairs, one 64-bit mask per chunk of pair slots
  unsigned long long live[IT];
#pragma unroll
  for (int k = 0; k < IT; k++) live[k] = __ballot(alive[k]);
  for (int iter = 0; iter < 4 * NP + 4; iter++) {
    // active nucleotides (count > 0) in ascending id order
    int nact = 0;
    unsigned long long am[ND];
#pragma unroll
    for (int q = 0; q < ND; q++) {
      am[q] = (q < nchunk_nodes) ? __ballot(ncnt[q] > 0) : 0ull;
      nact += __popcll(am[q]);
    }
    if (nact == 0) break;
    if (draw + 1 > w0 + 64u) {   // the pick draw must be in the window
      w0 = draw;
      wu = philox_u(seed, rk, step, root, w0 + (uint32_t)L);
      wl = log(wu);
      g2 = __ballot(wu > 0.5);
    }
    const int pick = bcast((int)ceil(wu * nact), (int)(draw - w0));   // 1..nact
    // the pick-th active nucleotide: its chunk (scalar prefix counts),
    // then the lane whose prefix count of active lanes is pick - 1
    int want = pick - 1;
    unsigned long long amq = am[0];
    int nidq = nid[0];
    {
      int acc = 0;
      bool found = false;
#pragma unroll
      for (int q = 0; q < ND; q++) {
        const int pc = __popcll(am[q]);
        const bool here = !found && acc + pc >= pick;
        want = here ? pick - acc - 1 : want;
        amq = here ? am[q] : amq;
        nidq = here ? nid[q] : nidq;
        found = found || here;
        acc += pc;
      }
    }
    const int ntd = bcast(nidq, __builtin_ctzll(__ballot(below(amq) == want) & amq));
    // formed pairs of ntd in list order
    unsigned long long sm[IT];
    int ns = 0;
#pragma unroll
    for (int k = 0; k < IT; k++) {
      sm[k] = (k < nchunk_items) ? ((__ballot(ia[k] == ntd) | __ballot(ib[k] == ntd)) & live[k]) : 0ull;
      ns += __popcll(sm[k]);
    }
    if (ns == 0) return false;
    if (ns > 64) ns = 64;
    // energy of pair slot `slot` (wave-uniform)
    auto ene_of = [&](int slot) {
      const int k = slot >> 6;
      double x = en[0];
#pragma unroll
      for (int kk = 1; kk < IT; kk++) x = (kk == k) ? en[kk] : x;
      return bcast(x, slot & 63);
    };
    int del;
    if (ns == 2) {
      // two pairs: the swap sweep leaves them reversed exactly when both
      // swap draws fall on the same side of one half
      if (draw + 4u > w0 + 64u) {
        w0 = draw;
        wu = philox_u(seed, rk, step, root, w0 + (uint32_t)L);
        wl = log(wu);
        g2 = __ballot(wu > 0.5);
      }
      const int o = (int)(draw - w0);
      int la = -1, lb = -1;
#pragma unroll
      for (int k = IT - 1; k >= 0; k--)
        if (sm[k]) la = 64 * k + __builtin_ctzll(sm[k]);
#pragma unroll
      for (int k = 0; k < IT; k++)
        if (sm[k]) lb = 64 * k + 63 - __builtin_clzll(sm[k]);
      const unsigned t = (unsigned)(g2 >> (o + 1)) & 3u;
      const bool flip = t == 0u || t == 3u;
      const int s0 = flip ? lb : la, s1 = flip ? la : lb;
      const double e0 = ene_of(s0), w = bcast(wl, o + 3);
      const int k1 = s1 >> 6;
      double e1v = en[0];
#pragma unroll
      for (int kk = 1; kk < IT; kk++) e1v = (kk == k1) ? en[kk] : e1v;
      const bool accd = (__ballot(w < (e1v - e0) * beta) >> (s1 & 63)) & 1ull;
      del = accd ? s1 : s0;
    } else if (ns <= 8) {
      // short sequence: kept in scalar registers, entry m in bits
      // 8m..8m+7 (slots < 256); the window holds the 2 ns draws
      uint64_t sp = 0;
      int m = 0;
#pragma unroll
      for (int k = 0; k < IT; k++)
        for (unsigned long long mk = sm[k]; mk; mk &= mk - 1) {
          sp |= (uint64_t)(64 * k + __builtin_ctzll(mk)) << (8 * m);
          m++;
        }
      if (draw + 2u * (uint32_t)ns > w0 + 64u) {
        w0 = draw;
        wu = philox_u(seed, rk, step, root, w0 + (uint32_t)L);
        wl = log(wu);
        g2 = __ballot(wu > 0.5);
      }
      const int o = (int)(draw - w0);
      // swap target of a shuffle draw held by this lane
      const int swv = (int)ceil(wu * ns) - 1;
      // reference shuffle: for i = 1..n swap(i, ceil(rnd*n))
      for (int i = 0; i < ns; i++) {
        const int sw = bcast(swv, o + 1 + i);
        const uint64_t x = ((sp >> (8 * i)) ^ (sp >> (8 * sw))) & 0xffull;   // swap entries i and sw
        sp ^= (x << (8 * i)) | (x << (8 * sw));
      }
      // Metropolis sweep on the energies of the sequence's pairs
      del = (int)(sp & 0xffull);
      double edel = ene_of(del);
      for (int i = 1; i < ns; i++) {
        const int si = (int)((sp >> (8 * i)) & 0xffull);
        const double ej = ene_of(si);
        if (bcast(wl, o + ns + i) < (ej - edel) * beta) { del = si; edel = ej; }
      }
    } else {
      // long sequence: entry m in lane m (through LDS)
#pragma unroll
      for (int k = 0, m0 = 0; k < IT; k++) {
        const int m = m0 + below(sm[k]);
        if (((sm[k] >> L) & 1ull) && m < 64) sseq[m] = 64 * k + L;
        m0 += __popcll(sm[k]);
      }
      wave_lds_fence();
      int seqv = L < ns ? sseq[L] : 0;
      // draws used: 1 (pick) + ns (shuffle) + ns - 1 (Metropolis) = 2 ns;
      // a window holding all of them starts at this iteration's pick draw
      const bool fits = 2 * ns <= 64;
      if (fits && draw + 2u * (uint32_t)ns > w0 + 64u) {
        w0 = draw;
        wu = philox_u(seed, rk, step, root, w0 + (uint32_t)L);
        wl = log(wu);
        g2 = __ballot(wu > 0.5);
      }
      const int o = (int)(draw - w0);
      // u and log u of the draw `k` places after the pick draw
      auto du = [&](int k) { return fits ? bcast(wu, o + k) : philox_u(seed, rk, step, root, draw + (uint32_t)k); };
      auto dlog = [&](int k) { return fits ? bcast(wl, o + k) : log(philox_u(seed, rk, step, root, draw + (uint32_t)k)); };
      // swap targets of the shuffle, lane i holding the i-th
      const int swl = fits ? (int)ceil(__shfl(wu, (o + 1 + L) & 63) * ns) - 1 : 0;
      for (int i = 0; i < ns; i++) {
        const int sw = fits ? bcast(swl, i) : (int)ceil(du(1 + i) * ns) - 1;
        const int ti = bcast(seqv, i), ts = bcast(seqv, sw);
        seqv = setlane(seqv, ts, i);
        seqv = setlane(seqv, ti, sw);
      }
      del = bcast(seqv, 0);
      double edel = ene_of(del);
      for (int i = 1; i < ns; i++) {
        const int si = bcast(seqv, i);
        const double ej = ene_of(si);
        if (dlog(ns + i) < (ej - edel) * beta) { del = si; edel = ej; }
      }
    }
    draw += 2u * (uint32_t)ns;
    // delete: pair slot del, decrement its two nucleotides
    const int dk = del >> 6, dl = del & 63;
    int xa = ia[0], xb = ib[0];
#pragma unroll
    for (int k = 1; k < IT; k++) {
      xa = (k == dk) ? ia[k] : xa;
      xb = (k == dk) ? ib[k] : xb;
    }
    const int da = bcast(xa, dl), db = bcast(xb, dl);
#pragma unroll
    for (int k = 0; k < IT; k++)
      if (k == dk) live[k] &= ~(1ull << dl);
#pragma unroll
    for (int q = 0; q < ND; q++) ncnt[q] -= (nid[q] == da) + (nid[q] == db);
  }
#pragma unroll
  for (int k = 0; k < IT; k++) alive[k] = (live[k] >> L) & 1ull;
  // write back the pairs deleted here
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int t = L + 64 * k;
    if (t < n && !alive[k]) D.bp_status[gb[k]] = 0;
  }
  return true;
}

// Resolve the component whose formed pairs are seg[0..n) (candidate ids,
// any order).  Must be called by all 64 lanes of a wave with uniform
// arguments; ws is mcw_lds(IT) ints of LDS private to the wave.  Returns
// false if the component exceeds the capacity of 64*IT pairs.  The result
// does not depend on IT (same draws, same decisions).  r indexes the per-replica
// arrays; rk keys the random stream.
// Layout: pair slot t (candidate-id order) lives in lane t % 64, chunk
// t / 64; nucleotide slot u (ascending id) in lane u % 64, chunk u / 64;
// the pair sequence of one iteration holds entry m in lane m.
template <int IT = MCW_ITEMS, int ND = MCW_NODES>
__device__ __forceinline__ bool mc_component_wave(const Params &P, const Ptrs &D, int r, int rk, int root,
                                                  const int *seg, int n, double beta, uint64_t seed,
                                                  int64_t step, int *ws) {
  constexpr int NONE = 0x7fffffff;
  constexpr int NP = 64 * IT, NE = 2 * NP;
  static_assert(NP <= 256, "pair slots are packed in 8 bits");
  const int L = lane_id();
  const int N = P.N;
  // uniform by contract; said so to the compiler, so that loop counters
  // and lane indices stay scalar
  n = __builtin_amdgcn_readfirstlane(n);
  root = __builtin_amdgcn_readfirstlane(root);
  if (n > NP) return false;
  const int nc = 2 * n;
  int *sraw = ws, *ssort = ws + NP, *send = ws + 2 * NP, *sflag = ws + 2 * NP + NE, *sseq = ws + 2 * NP + 2 * NE;
  // ---- pairs in candidate-id order (= reference list order): rank by
  // comparing against all ids (4 per LDS read), then scatter by rank
  const int n4 = (n + 3) & ~3;
  int raw[IT];
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int t = L + 64 * k;
    raw[k] = t < n ? seg[t] : NONE;
    if (t < n4) sraw[t] = raw[k];
  }
  wave_lds_fence();
  int rank[IT] = {};
  for (int j = 0; j < n4; j += 4) {
    const int4 v = *reinterpret_cast<const int4 *>(sraw + j);
#pragma unroll
    for (int k = 0; k < IT; k++)
      rank[k] += (v.x < raw[k]) + (v.y < raw[k]) + (v.z < raw[k]) + (v.w < raw[k]);
  }
#pragma unroll
  for (int k = 0; k < IT; k++)
    if (L + 64 * k < n) ssort[rank[k]] = raw[k];
  wave_lds_fence();
  int gb[IT], ia[IT], ib[IT];
  double en[IT];
  bool alive[IT];
  bool bad = false;
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int t = L + 64 * k;
    alive[k] = t < n;
    gb[k] = NONE;
    ia[k] = ib[k] = -1;
    en[k] = 0.0;
    if (alive[k]) {
      gb[k] = ssort[t];
      const int *mp = D.bp_mp + (size_t)gb[k] * 8;
      ia[k] = mp[0] - 1;
      ib[k] = mp[1] - 1;
      en[k] = D.ene_bp[gb[k]];
      bad = bad || ia[k] < 0 || ia[k] >= N || ib[k] < 0 || ib[k] >= N;
    }
  }
  // a candidate naming a bead outside the system means corrupted MC state:
  // report it (flags[7]) instead of indexing with it
  if (__ballot(bad)) {
    if (L == 0) D.flags[7] = 1;
    return true;
  }
  // ---- distinct nucleotides in ascending id order, with their counts.
  // Endpoint c = 2t (+1) is nucleotide a (b) of pair slot t.
  const int nc4 = (nc + 3) & ~3;
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int t = L + 64 * k;
    if (t < n) {
      send[2 * t] = ia[k];
      send[2 * t + 1] = ib[k];
    } else if (2 * t < nc4) {
      send[2 * t] = NONE;
      send[2 * t + 1] = NONE;
    }
  }
  wave_lds_fence();
  // first occurrence of each endpoint value
  int ev[ND];
  bool first[ND];
#pragma unroll
  for (int q = 0; q < ND; q++) {
    const int c = L + 64 * q;
    ev[q] = c < nc ? send[c] : NONE;
    first[q] = c < nc;
  }
  for (int j = 0; j < nc4; j += 4) {
    const int4 v = *reinterpret_cast<const int4 *>(send + j);
#pragma unroll
    for (int q = 0; q < ND; q++) {
      const int c = L + 64 * q;
      if ((j < c && v.x == ev[q]) || (j + 1 < c && v.y == ev[q]) || (j + 2 < c && v.z == ev[q]) ||
          (j + 3 < c && v.w == ev[q]))
        first[q] = false;
    }
  }
#pragma unroll
  for (int q = 0; q < ND; q++) {
    const int c = L + 64 * q;
    if (c < nc4) sflag[c] = first[q] ? ev[q] : NONE;   // first occurrences only
  }
  wave_lds_fence();
  // rank among distinct values, then scatter
  int nn = 0;
#pragma unroll
  for (int q = 0; q < ND; q++) nn += __popcll(__ballot(first[q]));
  int urank[ND] = {};
  for (int j = 0; j < nc4; j += 4) {
    const int4 v = *reinterpret_cast<const int4 *>(sflag + j);
#pragma unroll
    for (int q = 0; q < ND; q++)
      urank[q] += (v.x < ev[q]) + (v.y < ev[q]) + (v.z < ev[q]) + (v.w < ev[q]);
  }
#pragma unroll
  for (int q = 0; q < ND; q++)
    if (first[q]) send[urank[q]] = ev[q];   // endpoints no longer needed
  wave_lds_fence();
  int nid[ND], ncnt[ND];
#pragma unroll
  for (int q = 0; q < ND; q++) {
    const int u = L + 64 * q;
    nid[q] = u < nn ? send[u] : NONE;
    ncnt[q] = u < nn ? D.nt_cnt[(size_t)r * N + nid[q]] : 0;
  }
  // the sweep runs over as few nucleotide chunks as the component needs:
  // two for at most 128 distinct nucleotides (nearly all components), four
  // for at most 256
  if (ND > 2 && nn <= 128) {
    const int nid2[2] = {nid[0], nid[1]};
    int ncnt2[2] = {ncnt[0], ncnt[1]};
    return mcg_sweep<IT, 2>(P, D, rk, root, n, nn, beta, seed, step, sseq, gb, ia, ib, en, alive, nid2, ncnt2);
  }
  if (ND > 4 && nn <= 256) {
    const int nid4[4] = {nid[0], nid[1], nid[2], nid[3]};
    int ncnt4[4] = {ncnt[0], ncnt[1], ncnt[2], ncnt[3]};
    return mcg_sweep<IT, 4>(P, D, rk, root, n, nn, beta, seed, step, sseq, gb, ia, ib, en, alive, nid4, ncnt4);
  }
  return mcg_sweep<IT, ND>(P, D, rk, root, n, nn, beta, seed, step, sseq, gb, ia, ib, en, alive, nid, ncnt);
}

// ---------------------------------------------------------------------
// The same sweep for a component of at most 64 pairs (the usual size),
// with one pair per lane and the bookkeeping in scalar registers: the live
// pairs are a 64-bit mask, the nucleotides two slots per lane, the sequence
// of up to 8 pairs held as a 4-bit permutation of their ranks.  A single
// wavefront runs this serially, so its cost is the latency of its chain of
// lane reads and ballots (each about 25-30 cycles on gfx950); the common
// two-pair case resolves its swap sweep from one ballot of the block of
// draws taken at each refill.  Draws and decisions are those of
// mc_component_wave.  Returns false (having written nothing) if the
// component is larger, a nucleotide has more than 8 formed pairs, or the
// state is inconsistent; the caller then runs the general form from the
// start, which reproduces the same result.
__device__ __forceinline__ bool mc_component_wave64(const Params &P, const Ptrs &D, int r, int rk, int root,
                                                    const int *seg, int n, double beta, uint64_t seed, int64_t step,
                                                    int *ws) {
  constexpr int NONE = 0x7fffffff;
  const int L = lane_id();
  const int N = P.N;
  n = __builtin_amdgcn_readfirstlane(n);
  root = __builtin_amdgcn_readfirstlane(root);
  if (n > 64) return false;
  int *sraw = ws, *ssort = ws + 64, *send = ws + 128, *sflag = ws + 256;   // 384 ints
  // ---- pairs in candidate-id order: rank by comparison, scatter
  const int n4 = (n + 3) & ~3;
  const int raw = L < n ? seg[L] : NONE;
  if (L < n4) sraw[L] = raw;
  wave_lds_fence();
  int rank = 0;
  for (int j = 0; j < n4; j += 4) {
    const int4 v = *reinterpret_cast<const int4 *>(sraw + j);
    rank += (v.x < raw) + (v.y < raw) + (v.z < raw) + (v.w < raw);
  }
  if (L < n) ssort[rank] = raw;
  wave_lds_fence();
  int gb = NONE, ia = -1, ib = -1;
  double en = 0.0;
  bool bad = false;
  if (L < n) {
    gb = ssort[L];
    const int *mp = D.bp_mp + (size_t)gb * 8;
    ia = mp[0] - 1;
    ib = mp[1] - 1;
    en = D.ene_bp[gb];
    bad = ia < 0 || ia >= N || ib < 0 || ib >= N;
  }
  if (__ballot(bad)) return false;
  // ---- distinct nucleotides in ascending id order (slots L and 64 + L)
  const int nc = 2 * n, nc4 = (nc + 3) & ~3;
  if (L < n) {
    send[2 * L] = ia;
    send[2 * L + 1] = ib;
  } else if (2 * L < nc4) {
    send[2 * L] = NONE;
    send[2 * L + 1] = NONE;
  }
  wave_lds_fence();
  const int c1 = 64 + L;
  const int ev0 = L < nc ? send[L] : NONE, ev1 = c1 < nc ? send[c1] : NONE;
  bool f0 = L < nc, f1 = c1 < nc;
  for (int j = 0; j < nc4; j += 4) {
    const int4 v = *reinterpret_cast<const int4 *>(send + j);
    if ((j < L && v.x == ev0) || (j + 1 < L && v.y == ev0) || (j + 2 < L && v.z == ev0) || (j + 3 < L && v.w == ev0))
      f0 = false;
    if ((j < c1 && v.x == ev1) || (j + 1 < c1 && v.y == ev1) || (j + 2 < c1 && v.z == ev1) ||
        (j + 3 < c1 && v.w == ev1))
      f1 = false;
  }
  if (L < nc4) sflag[L] = f0 ? ev0 : NONE;
  if (c1 < nc4) sflag[c1] = f1 ? ev1 : NONE;
  wave_lds_fence();
  const int nn = __popcll(__ballot(f0)) + __popcll(__ballot(f1));
  int u0 = 0, u1 = 0;
  for (int j = 0; j < nc4; j += 4) {
    const int4 v = *reinterpret_cast<const int4 *>(sflag + j);
    u0 += (v.x < ev0) + (v.y < ev0) + (v.z < ev0) + (v.w < ev0);
    u1 += (v.x < ev1) + (v.y < ev1) + (v.z < ev1) + (v.w < ev1);
  }
  if (f0) send[u0] = ev0;
  if (f1) send[u1] = ev1;
  wave_lds_fence();
  const int nid0 = L < nn ? send[L] : NONE, nid1 = c1 < nn ? send[c1] : NONE;
  int cnt0 = L < nn ? D.nt_cnt[(size_t)r * N + nid0] : 0;
  int cnt1 = c1 < nn ? D.nt_cnt[(size_t)r * N + nid1] : 0;
  // ---- the sweep
  unsigned long long live = __ballot(L < n);
  uint32_t draw = 0, w0 = 0;
  double wu = philox_u(seed, rk, step, root, (uint32_t)L), wl = log(wu);
  // draws above one half: the swap target of a two-pair sequence is 1 there
  unsigned long long g2 = __ballot(wu > 0.5);
  for (int iter = 0; iter < 4 * 64 + 4; iter++) {
    const unsigned long long a0 = __ballot(cnt0 > 0), a1 = __ballot(cnt1 > 0);
    const int p0 = __popcll(a0), nact = p0 + __popcll(a1);
    if (nact == 0) break;
    if (draw + 1 > w0 + 64u) {
      w0 = draw;
      wu = philox_u(seed, rk, step, root, w0 + (uint32_t)L);
      wl = log(wu);
      g2 = __ballot(wu > 0.5);
    }
    const int pick = bcast((int)ceil(wu * nact), (int)(draw - w0));   // 1..nact
    const bool upper = pick > p0;
    const unsigned long long am = upper ? a1 : a0;
    const int want = upper ? pick - p0 - 1 : pick - 1;
    const int ln = __builtin_ctzll(__ballot(below(am) == want) & am);
    const int ntd = bcast(upper ? nid1 : nid0, ln);
    const unsigned long long sm = (__ballot(ia == ntd) | __ballot(ib == ntd)) & live;
    const int ns = __popcll(sm);
    if (ns == 0 || ns > 8) return false;
    if (draw + 2u * (uint32_t)ns > w0 + 64u) {
      w0 = draw;
      wu = philox_u(seed, rk, step, root, w0 + (uint32_t)L);
      wl = log(wu);
      g2 = __ballot(wu > 0.5);
    }
    const int o = (int)(draw - w0);
    int del;
    if (ns == 2) {
      // two pairs: the swap sweep leaves them reversed exactly when both
      // swap draws fall on the same side of one half
      const int la = __builtin_ctzll(sm), lb = 63 - __builtin_clzll(sm);
      const unsigned t = (unsigned)(g2 >> (o + 1)) & 3u;
      const bool flip = t == 0u || t == 3u;
      const int s0 = flip ? lb : la, s1 = flip ? la : lb;
      const double e0 = bcast(en, s0), w = bcast(wl, o + 3);
      const bool acc = (__ballot(w < (en - e0) * beta) >> s1) & 1ull;
      del = acc ? s1 : s0;
    } else {
      const int swv = (int)ceil(wu * ns) - 1;
      // ntd's pairs in list order are the set bits of sm in lane order (rank
      // 0..ns-1); perm holds the rank at each sequence slot, 4 bits a slot
      uint32_t perm = 0x76543210u;
      for (int i = 0; i < ns; i++) {
        const int sw = bcast(swv, o + 1 + i);
        const uint32_t x = ((perm >> (4 * i)) ^ (perm >> (4 * sw))) & 15u;
        perm ^= (x << (4 * i)) | (x << (4 * sw));
      }
      const int rkl = below(sm);
      del = __builtin_ctzll(__ballot(rkl == (int)(perm & 15u)) & sm);
      double edel = bcast(en, del);
      for (int i = 1; i < ns; i++) {
        const int si = __builtin_ctzll(__ballot(rkl == (int)((perm >> (4 * i)) & 15u)) & sm);
        const double ej = bcast(en, si);
        if (bcast(wl, o + ns + i) < (ej - edel) * beta) { del = si; edel = ej; }
      }
    }
    draw += 2u * (uint32_t)ns;
    live &= ~(1ull << del);
    const int da = bcast(ia, del), db = bcast(ib, del);
    cnt0 -= (nid0 == da) + (nid0 == db);
    cnt1 -= (nid1 == da) + (nid1 == db);
  }
  if (L < n && !((live >> L) & 1ull)) D.bp_status[gb] = 0;
  return true;
}

// The wavefront forms by size: 64 pairs or fewer in the scalar form, up
// to 64 * MCW_ITEMS in the general one (also the fallback of the scalar
// form).  ws holds MCW_LDS ints.
__device__ __forceinline__ bool mc_component_wave_any(const Params &P, const Ptrs &D, int r, int rk, int root,
                                                      const int *seg, int n, double beta, uint64_t seed,
                                                      int64_t step, int *ws) {
  if (n <= 64 && mc_component_wave64(P, D, r, rk, root, seg, n, beta, seed, step, ws)) return true;
  return mc_component_wave(P, D, r, rk, root, seg, n, beta, seed, step, ws);
}

// ---------------------------------------------------------------------
// The same algorithm for a component of at most MCL_PAIRS pairs, run by
// one lane with everything in registers: arrays are read through unrolled
// selects and the pair sequence is packed in 4-bit fields, so no per-lane
// array is indexed dynamically.  The draws (seed, rk, step, root, 0, 1, ...)
// and every decision are those of mc_component_wave, so the two give
// identical results; many small components resolve in parallel, one per
// lane.  Returns false if the component is too large or inconsistent.
constexpr int MCL_PAIRS = 4;

template <int M>
__device__ __forceinline__ int sel_i(const int (&a)[M], int i) {
  int v = a[0];
#pragma unroll
  for (int k = 1; k < M; k++) v = (i == k) ? a[k] : v;
  return v;
}
template <int M>
__device__ __forceinline__ double sel_d(const double (&a)[M], int i) {
  double v = a[0];
#pragma unroll
  for (int k = 1; k < M; k++) v = (i == k) ? a[k] : v;
  return v;
}

__device__ __forceinline__ bool mc_component_lane(const Ptrs &D, int N, int r, int rk, int root, const int *seg,
                                                  int n, double beta, uint64_t seed, int64_t step) {
  constexpr int M = MCL_PAIRS, M2 = 2 * MCL_PAIRS;
  constexpr int NONE = 0x7fffffff;
  if (n > M) return false;
  // candidate ids in ascending order (= list order)
  int raw[M], g[M];
#pragma unroll
  for (int j = 0; j < M; j++) raw[j] = j < n ? seg[j] : NONE;
#pragma unroll
  for (int p = 0; p < M; p++) g[p] = NONE;
#pragma unroll
  for (int i = 0; i < M; i++) {
    int rank = 0;
#pragma unroll
    for (int j = 0; j < M; j++) rank += raw[j] < raw[i] ? 1 : 0;
#pragma unroll
    for (int p = 0; p < M; p++)
      if (i < n && rank == p) g[p] = raw[i];
  }
  int pa[M], pb[M];
  double en[M];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < M; j++) {
    pa[j] = pb[j] = -1;
    en[j] = 0.0;
    if (j < n) {
      const int *mp = D.bp_mp + (size_t)g[j] * 8;
      pa[j] = mp[0] - 1;
      pb[j] = mp[1] - 1;
      en[j] = D.ene_bp[g[j]];
      bad = bad || pa[j] < 0 || pa[j] >= N || pb[j] < 0 || pb[j] >= N;
    }
  }
  if (bad) {
    D.flags[7] = 1;
    return true;
  }
  // distinct nucleotides in ascending id order, with their counts
  int e[M2];
#pragma unroll
  for (int j = 0; j < M; j++) {
    e[2 * j] = j < n ? pa[j] : NONE;
    e[2 * j + 1] = j < n ? pb[j] : NONE;
  }
  bool first[M2];
#pragma unroll
  for (int k = 0; k < M2; k++) {
    bool f = e[k] != NONE;
#pragma unroll
    for (int j = 0; j < M2; j++)
      if (j < k && e[j] == e[k]) f = false;
    first[k] = f;
  }
  int node[M2], cnt[M2];
#pragma unroll
  for (int q = 0; q < M2; q++) node[q] = NONE;
#pragma unroll
  for (int k = 0; k < M2; k++) {
    int ur = 0;
#pragma unroll
    for (int j = 0; j < M2; j++) ur += (first[j] && e[j] < e[k]) ? 1 : 0;
#pragma unroll
    for (int q = 0; q < M2; q++)
      if (first[k] && ur == q) node[q] = e[k];
  }
#pragma unroll
  for (int q = 0; q < M2; q++) cnt[q] = node[q] != NONE ? D.nt_cnt[(size_t)r * N + node[q]] : 0;
  // the Monte Carlo loop
  uint32_t alive = (n >= 32) ? 0xffffffffu : ((1u << n) - 1u);
  uint32_t draw = 0;
  for (int iter = 0; iter < 4 * M + 4; iter++) {
    int nact = 0;
#pragma unroll
    for (int q = 0; q < M2; q++) nact += cnt[q] > 0 ? 1 : 0;
    if (nact == 0) break;
    const int pick = (int)ceil(philox_u(seed, rk, step, root, draw++) * nact);   // 1..nact
    int ntd = -1, acc = 0;
#pragma unroll
    for (int q = 0; q < M2; q++)
      if (cnt[q] > 0) {
        acc++;
        if (acc == pick) ntd = node[q];
      }
    // formed pairs of ntd in list order, packed 4 bits each
    uint32_t seqp = 0;
    int ns = 0;
#pragma unroll
    for (int j = 0; j < M; j++)
      if (((alive >> j) & 1u) && (pa[j] == ntd || pb[j] == ntd)) {
        seqp |= (uint32_t)j << (4 * ns);
        ns++;
      }
    if (ns == 0) return false;
    // reference shuffle: for i = 1..n swap(i, ceil(rnd*n))
    for (int i = 0; i < ns; i++) {
      const int sw = (int)ceil(philox_u(seed, rk, step, root, draw++) * ns) - 1;
      const uint32_t ti = (seqp >> (4 * i)) & 15u, ts = (seqp >> (4 * sw)) & 15u;
      seqp = (seqp & ~(15u << (4 * i))) | (ts << (4 * i));
      seqp = (seqp & ~(15u << (4 * sw))) | (ti << (4 * sw));
    }
    // Metropolis sweep on the energies
    int del = (int)(seqp & 15u);
    double edel = sel_d(en, del);
    for (int i = 1; i < ns; i++) {
      const int j = (int)((seqp >> (4 * i)) & 15u);
      const double ej = sel_d(en, j);
      if (log(philox_u(seed, rk, step, root, draw++)) < (ej - edel) * beta) {
        del = j;
        edel = ej;
      }
    }
    alive &= ~(1u << del);
    const int da = sel_i(pa, del), db = sel_i(pb, del);
#pragma unroll
    for (int q = 0; q < M2; q++)
      if (node[q] == da || node[q] == db) cnt[q] -= 1;
  }
#pragma unroll
  for (int j = 0; j < M; j++)
    if (j < n && !((alive >> j) & 1u)) D.bp_status[g[j]] = 0;
  return true;
}

}  // namespace sis
