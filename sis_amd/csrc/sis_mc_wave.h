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

constexpr int MCW_ITEMS = 4;   // pairs per lane   -> 256 pairs per component
constexpr int MCW_NODES = 8;   // nodes per lane   -> 512 nucleotides

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double philox_u(uint64_t seed, int r, int64_t step, int root, uint32_t draw) {
  U4 o = philox(U4{(uint32_t)step, (uint32_t)(step >> 32) ^ 0x4d43u, (uint32_t)root, draw},
                (uint32_t)seed, (uint32_t)(seed >> 32) ^ (0x5bd1e995u * (uint32_t)(r + 1)));
  return u01(o.x, o.y);
}

// Resolve the component whose formed pairs are seg[0..n) (candidate ids,
// any order).  Must be called by all 64 lanes of a wave with uniform
// arguments.  Returns false if the component exceeds the wave capacity.
// r indexes the per-replica arrays; rk keys the random stream.
__device__ __noinline__ bool mc_component_wave(const Params &P, const Ptrs &D, int r, int rk, int root,
                                               const int *seg, int n, double beta, uint64_t seed,
                                               int64_t step, int *seq) {
  const int L = lane_id();
  const int N = P.N;
  if (n > 64 * MCW_ITEMS) return false;
  // ---- load pairs, order them by candidate id (= reference list order)
  int gb[MCW_ITEMS], ia[MCW_ITEMS], ib[MCW_ITEMS];
  double en[MCW_ITEMS];
  bool alive[MCW_ITEMS];
  int raw[MCW_ITEMS];
#pragma unroll
  for (int k = 0; k < MCW_ITEMS; k++) {
    int t = L + 64 * k;
    raw[k] = t < n ? seg[t] : 0x7fffffff;
  }
  // rank of each raw id among all n ids
  int rank[MCW_ITEMS];
#pragma unroll
  for (int k = 0; k < MCW_ITEMS; k++) rank[k] = 0;
  for (int c = 0; c < MCW_ITEMS; c++) {
    if (64 * c >= n) break;
    for (int s = 0; s < 64; s++) {
      int v = __shfl(raw[c], s);
#pragma unroll
      for (int k = 0; k < MCW_ITEMS; k++) rank[k] += (v < raw[k]) ? 1 : 0;
    }
  }
  // sorted[t] = raw id with rank t: gather by searching ranks
#pragma unroll
  for (int k = 0; k < MCW_ITEMS; k++) gb[k] = 0x7fffffff;
  for (int c = 0; c < MCW_ITEMS; c++) {
    if (64 * c >= n) break;
    for (int s = 0; s < 64; s++) {
      int v = __shfl(raw[c], s), rk = __shfl(rank[c], s);
      if (64 * c + s >= n) break;
#pragma unroll
      for (int k = 0; k < MCW_ITEMS; k++)
        if (rk == L + 64 * k) gb[k] = v;
    }
  }
#pragma unroll
  for (int k = 0; k < MCW_ITEMS; k++) {
    int t = L + 64 * k;
    alive[k] = t < n;
    if (alive[k]) {
      const int *mp = D.bp_mp + (size_t)gb[k] * 8;
      ia[k] = mp[0] - 1;
      ib[k] = mp[1] - 1;
      en[k] = D.ene_bp[gb[k]];
    } else {
      ia[k] = ib[k] = -1;
      en[k] = 0.0;
    }
  }
  // a candidate naming a bead outside the system means corrupted MC state:
  // report it (flags[7]) instead of indexing with it
  {
    bool bad = false;
#pragma unroll
    for (int k = 0; k < MCW_ITEMS; k++)
      bad = bad || (alive[k] && (ia[k] < 0 || ia[k] >= N || ib[k] < 0 || ib[k] >= N));
    if (__ballot(bad)) {
      if (L == 0) D.flags[7] = 1;
      return true;
    }
  }
  // ---- unique nucleotides in ascending id order, with their counts
  // candidate c = 2*t (+1) is endpoint a (b) of pair t
  int nid[MCW_NODES];
  int ncnt[MCW_NODES];
  const int nc = 2 * n;
  if (nc > 64 * MCW_NODES) return false;
  int cand[MCW_NODES];
#pragma unroll
  for (int q = 0; q < MCW_NODES; q++) {
    int c = L + 64 * q;
    int t = (c < nc) ? (c >> 1) : 0;
    int kk = t / 64, ll = t % 64;
    int a = 0, b = 0;
#pragma unroll
    for (int k = 0; k < MCW_ITEMS; k++) {
      int aa = __shfl(ia[k], ll), bb = __shfl(ib[k], ll);   // all lanes active
      if (k == kk) { a = aa; b = bb; }
    }
    cand[q] = (c < nc) ? ((c & 1) ? b : a) : 0x7fffffff;
  }
  // first occurrence flag and rank among unique ids
  bool first[MCW_NODES];
#pragma unroll
  for (int q = 0; q < MCW_NODES; q++) first[q] = (L + 64 * q) < nc;
  for (int c = 0; c < MCW_NODES; c++) {
    if (64 * c >= nc) break;
    for (int s = 0; s < 64; s++) {
      int v = __shfl(cand[c], s);
      int pos = 64 * c + s;
#pragma unroll
      for (int q = 0; q < MCW_NODES; q++)
        if (pos < L + 64 * q && v == cand[q]) first[q] = false;
    }
  }
  int urank[MCW_NODES];
#pragma unroll
  for (int q = 0; q < MCW_NODES; q++) urank[q] = 0;
  for (int c = 0; c < MCW_NODES; c++) {
    if (64 * c >= nc) break;
    for (int s = 0; s < 64; s++) {
      int v = __shfl(cand[c], s);
      int f = __shfl((int)first[c], s);
#pragma unroll
      for (int q = 0; q < MCW_NODES; q++)
        if (f && v < cand[q]) urank[q]++;
    }
  }
  // place unique ids at their rank: node slot u (lane u%64, chunk u/64)
#pragma unroll
  for (int q = 0; q < MCW_NODES; q++) nid[q] = 0x7fffffff;
  for (int c = 0; c < MCW_NODES; c++) {
    if (64 * c >= nc) break;
    for (int s = 0; s < 64; s++) {
      int v = __shfl(cand[c], s), f = __shfl((int)first[c], s), rk = __shfl(urank[c], s);
      if (f) {
#pragma unroll
        for (int q = 0; q < MCW_NODES; q++)
          if (rk == L + 64 * q) nid[q] = v;
      }
    }
  }
  int nn = 0;
  {
    int cntu = 0;
#pragma unroll
    for (int q = 0; q < MCW_NODES; q++) cntu += (nid[q] != 0x7fffffff) ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) cntu += __shfl_xor(cntu, off);
    nn = cntu;
  }
#pragma unroll
  for (int q = 0; q < MCW_NODES; q++)
    ncnt[q] = (nid[q] != 0x7fffffff) ? D.nt_cnt[(size_t)r * N + nid[q]] : 0;
  // ---- the Monte Carlo loop
  uint32_t draw = 0;
  const int nchunk_nodes = (nn + 63) / 64;
  const int nchunk_items = (n + 63) / 64;
  for (int iter = 0; iter < 4 * 64 * MCW_ITEMS + 4; iter++) {
    // active nucleotides (count > 0) in ascending id order
    int nact = 0;
    unsigned long long am[MCW_NODES];
#pragma unroll
    for (int q = 0; q < MCW_NODES; q++) {
      am[q] = (q < nchunk_nodes) ? __ballot(ncnt[q] > 0) : 0ull;
      nact += __popcll(am[q]);
    }
    if (nact == 0) break;
    double u = philox_u(seed, rk, step, root, draw++);
    int pick = (int)ceil(u * nact);   // 1..nact
    int ntd = -1;
    {
      int acc = 0;
#pragma unroll
      for (int q = 0; q < MCW_NODES; q++) {
        int pc = __popcll(am[q]);
        if (ntd < 0 && acc + pc >= pick) {
          int want = pick - acc;   // want-th set bit of am[q]
          unsigned long long m = am[q];
          for (int w = 1; w < want; w++) m &= m - 1;
          int lanebit = __ffsll((long long)m) - 1;
          ntd = __shfl(nid[q], lanebit);
        }
        acc += pc;
      }
    }
    // formed pairs of ntd in list order
    unsigned long long sm[MCW_ITEMS];
    int ns = 0;
#pragma unroll
    for (int k = 0; k < MCW_ITEMS; k++) {
      sm[k] = (k < nchunk_items) ? __ballot(alive[k] && (ia[k] == ntd || ib[k] == ntd)) : 0ull;
      ns += __popcll(sm[k]);
    }
    if (ns == 0) return false;
    // sequence of pair slots (slot = 64*k + lane); every lane holds the
    // same values and writes the same LDS words
    int m = 0;
#pragma unroll
    for (int k = 0; k < MCW_ITEMS; k++) {
      unsigned long long b = sm[k];
      while (b && m < 64) {
        int l = __ffsll((long long)b) - 1;
        seq[m++] = 64 * k + l;
        b &= b - 1;
      }
    }
    if (ns > 64) ns = 64;
    // reference shuffle: for i = 1..n swap(i, ceil(rnd*n))
    for (int i = 0; i < ns; i++) {
      double v = philox_u(seed, rk, step, root, draw++);
      int sw = (int)ceil(v * ns) - 1;
      int t = seq[i];
      seq[i] = seq[sw];
      seq[sw] = t;
    }
    // energies of the sequence's pairs
    auto ene_of = [&](int slot) {
      int k = slot / 64, l = slot % 64;
      double e = 0.0;
#pragma unroll
      for (int kk = 0; kk < MCW_ITEMS; kk++) {
        double x = __shfl(en[kk], l);
        if (kk == k) e = x;
      }
      return e;
    };
    int del = seq[0];
    double edel = ene_of(del);
    for (int i = 1; i < ns; i++) {
      double ej = ene_of(seq[i]);
      double ratio = exp((ej - edel) * beta);
      double v = philox_u(seed, rk, step, root, draw++);
      if (v < ratio) { del = seq[i]; edel = ej; }
    }
    // delete: pair slot del, decrement its two nucleotides
    int dk = del / 64, dl = del % 64;
    int da = -1, db = -1;
#pragma unroll
    for (int k = 0; k < MCW_ITEMS; k++) {
      int aa = __shfl(ia[k], dl), bb = __shfl(ib[k], dl);
      if (k == dk) { da = aa; db = bb; }
    }
#pragma unroll
    for (int k = 0; k < MCW_ITEMS; k++)
      if (k == dk && L == dl) alive[k] = false;
#pragma unroll
    for (int q = 0; q < MCW_NODES; q++)
      if (nid[q] == da || nid[q] == db) ncnt[q] -= 1;
  }
  // write back the pairs deleted here
#pragma unroll
  for (int k = 0; k < MCW_ITEMS; k++) {
    int t = L + 64 * k;
    if (t < n && !alive[k]) D.bp_status[gb[k]] = 0;
  }
  return true;
}

}  // namespace sis
