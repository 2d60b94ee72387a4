// Persistent step kernel for one large system (mode 2).  One cooperative
// launch runs job_md steps back to back; phases are separated by a grid
// barrier:
//   A  base-pair evaluation per candidate (+ union of formed pairs)
//   B  formed pairs bucketed by component root          (MC steps only)
//   C  one wavefront per component with excess: limit MC (MC steps only)
//   D  owner-computes force + GJF-2GJ per bead (lane groups), skin top-2
// then every block reads the per-block top-2 displacements and makes the
// same rebuild decision (src/job_md.F90:376-395); a rebuild ends the launch
// and the host rebuilds the lists.
#pragma once
#include "sis_kernels.h"

namespace sis {

struct PersArgs {
  int64_t nsteps, step0;
  uint64_t seed;
  unsigned long long *bar;  // monotonic barrier counter, zeroed before launch
  int *seg_n, *seg, *has_exc;
  int *done;                // [0] steps completed, [1] rebuild requested, [2] buffer parity
  double4 *pos_b;           // second position buffer
  int nbp_total, force_mc_first;
};

// Grid-wide barrier on a monotonic counter.  Agent-scope release before
// arriving and acquire after leaving make the phase's global writes of
// every block visible to every block (across XCDs).
__device__ __forceinline__ void grid_barrier(unsigned long long *bar, unsigned long long &target,
                                             unsigned nblocks) {
  __syncthreads();
  target += nblocks;
  if (threadIdx.x == 0) {
    __atomic_thread_fence(__ATOMIC_RELEASE);
    __hip_atomic_fetch_add(bar, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target)
      __builtin_amdgcn_s_sleep(1);
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
  }
  __syncthreads();
}

// one candidate of phase A (src/force_bp_limit_triplet.F90:85-256);
// returns 1 if formed
template <typename T>
__device__ __forceinline__ int bp_eval_one(const Params &P, const Ptrs &D, int gb, int mc) {
  const int *mp = D.bp_mp + (size_t)gb * 8;
  const int r = mp[7];
  const int N = P.N;
  D.bp_status[gb] = 0;
  D.ene_bp[gb] = 0.0;
  if (!mc && P.nstep_bp_mc > 0 && !D.bp_status_mc[gb]) return 0;
  double tK = D.tK[r];
  double dG = D.bp_coef[2 * (size_t)gb] - tK * D.bp_coef[2 * (size_t)gb + 1];
  if (dG >= 0.0) return 0;
  const double *box = D.box + 3 * r;
  const double4 *x = D.pos + (size_t)r * N;
  double4 p1 = x[mp[0] - 1], p2 = x[mp[1] - 1];
  V3<double> v12 = pvec(P, box, p1, p2), v13 = pvec(P, box, p1, x[mp[2] - 1]),
             v15 = pvec(P, box, p1, x[mp[4] - 1]), v42 = pvec(P, box, x[mp[3] - 1], p2),
             v62 = pvec(P, box, x[mp[5] - 1], p2);
  const BpPara &bp = P.bpp[mp[6]];
  T u;
  V3<T> fb[6];
  if (!bp_geom<T>(bp, v12, v13, v15, v42, v62, u, fb)) return 0;
  double e = bp.U0 * dG * exp(-(double)u);
  if (e > P.bp_cutoff_energy) return 0;
  D.bp_status[gb] = 1;
  D.ene_bp[gb] = e;
  double *fo = D.for_bp + (size_t)gb * 18;
  for (int q = 0; q < 6; q++) {
    fo[3 * q] = e * (double)fb[q].x;
    fo[3 * q + 1] = e * (double)fb[q].y;
    fo[3 * q + 2] = e * (double)fb[q].z;
  }
  atomicAdd(&D.nt_cnt[(size_t)r * N + mp[0] - 1], 1);
  atomicAdd(&D.nt_cnt[(size_t)r * N + mp[1] - 1], 1);
  return 1;
}

template <typename T, int LPB, int B>
__global__ void __launch_bounds__(B) k_persistent(Params P, Ptrs D, PersArgs A) {
  __shared__ double s1[B / LPB], s2[B / LPB];
  __shared__ int sseq[B];   // one 64-entry sequence per wave
  __shared__ int s_exit;
  const int N = P.N;
  const unsigned G = gridDim.x;
  const int gtid = blockIdx.x * B + threadIdx.x, gthreads = G * B;
  unsigned long long target = 0;
  const double BOLTZ = 1.380649e-23 * ((1.0 / 4184.0) * 6.02214076e23);
  const double beta = 1.0 / (BOLTZ * D.tK[0]);
  double4 *pos_a = D.pos, *pos_b = A.pos_b;
  int64_t completed = 0;
  int rebuild = 0;
  for (int64_t s = 0; s < A.nsteps; s++) {
    const int64_t step = A.step0 + s + 1;
    const int mc = P.nstep_bp_mc > 0 && ((step % P.nstep_bp_mc) == 0 || (s == 0 && A.force_mc_first));
    D.pos = pos_a;
    D.pos_out = pos_b;
    // ---- A: base pairs
    for (int gb = gtid; gb < A.nbp_total; gb += gthreads) {
      if (bp_eval_one<T>(P, D, gb, mc) && mc) {
        const int *mp = D.bp_mp + (size_t)gb * 8;
        int a = mp[0] - 1, b = mp[1] - 1;
        while (true) {
          a = uf_find(D.comp, a);
          b = uf_find(D.comp, b);
          if (a == b) break;
          if (a > b) { int t = a; a = b; b = t; }
          if (atomicCAS(&D.comp[b], b, a) == b) break;
        }
      }
    }
    grid_barrier(A.bar, target, G);
    if (mc) {
      // ---- B: bucket formed pairs per component root
      for (int gb = gtid; gb < A.nbp_total; gb += gthreads) {
        if (!D.bp_status[gb]) continue;
        const int *mp = D.bp_mp + (size_t)gb * 8;
        int a = mp[0] - 1, b = mp[1] - 1;
        int root = uf_find(D.comp, a);
        int slot = atomicAdd(&A.seg_n[root], 1);
        if (slot < SEGCAP) A.seg[(size_t)root * SEGCAP + slot] = gb;
        if (D.nt_cnt[a] > 0 || D.nt_cnt[b] > 0) A.has_exc[root] = 1;
      }
      grid_barrier(A.bar, target, G);
      // ---- C: one wavefront per component with excess
      {
        const int wave = gtid / 64, nwaves = gthreads / 64;
        int *seq = sseq + (threadIdx.x / 64) * 64;
        for (int root = wave; root < N; root += nwaves) {
          if (!A.has_exc[root]) continue;
          int n = A.seg_n[root];
          if (n > SEGCAP) {
            if ((threadIdx.x & 63) == 0) D.flags[5] = 1;
            continue;
          }
          int *sg = A.seg + (size_t)root * SEGCAP;
          if (!mc_component_wave(P, D, 0, root, sg, n, beta, A.seed, step, seq)) {
            if ((threadIdx.x & 63) == 0) mc_component_serial(P, D, 0, root, sg, n, beta, A.seed, step);
          }
        }
      }
      grid_barrier(A.bar, target, G);
    }
    // ---- D: forces and the update; reset per-step MC state
    double m1 = 0.0, m2 = 0.0;
    {
      const int grp = gtid / LPB, l = gtid % LPB, ngrp = gthreads / LPB;
      for (int a0 = 0; a0 < N; a0 += ngrp) {
        int a = a0 + grp;
        V3<double> F{0.0, 0.0, 0.0};
        if (a < N) F = bead_force_part<T>(P, D, 0, a, l, LPB);
        F = group_sum<LPB>(F);
        if (a < N && l == 0) {
          D.nt_cnt[a] = -P.max_bp_per_nt;
          D.comp[a] = a;
          A.seg_n[a] = 0;
          A.has_exc[a] = 0;
          double d2 = gjf_update(P, D, 0, a, F, 0, A.seed, step);
          if (d2 > m1) { m2 = m1; m1 = d2; }
          else if (d2 > m2) m2 = d2;
        }
      }
    }
    // block top-2 of the displacement
    {
      const int grp = threadIdx.x / LPB, l = threadIdx.x % LPB;
      if (l == 0) { s1[grp] = m1; s2[grp] = m2; }
      __syncthreads();
      for (int off = (B / LPB) / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
          double a1 = s1[threadIdx.x], a2 = s2[threadIdx.x], b1 = s1[threadIdx.x + off], b2 = s2[threadIdx.x + off];
          s1[threadIdx.x] = a1 > b1 ? a1 : b1;
          s2[threadIdx.x] = a1 > b1 ? (a2 > b1 ? a2 : b1) : (a1 > b2 ? a1 : b2);
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        D.blk_top2[2 * blockIdx.x] = s1[0];
        D.blk_top2[2 * blockIdx.x + 1] = s2[0];
      }
    }
    grid_barrier(A.bar, target, G);
    // swap position buffers: the updated ones are current from now on
    { double4 *t = pos_a; pos_a = pos_b; pos_b = t; }
    completed = s + 1;
    // same decision in every block
    if (threadIdx.x == 0) {
      double a1 = 0.0, a2 = 0.0;
      for (unsigned b = 0; b < G; b++) {
        double b1 = D.blk_top2[2 * b], b2 = D.blk_top2[2 * b + 1];
        double n1 = a1 > b1 ? a1 : b1;
        double n2 = a1 > b1 ? (a2 > b1 ? a2 : b1) : (a1 > b2 ? a1 : b2);
        a1 = n1;
        a2 = n2;
      }
      s_exit = (sqrt(a1) + sqrt(a2) > P.nl_margin) ? 1 : 0;
    }
    __syncthreads();
    rebuild = s_exit;
    // blk_top2 is rewritten in the next phase D only after two more barriers
    if (rebuild) break;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.done[0] = (int)completed;
    A.done[1] = rebuild;
    A.done[2] = (pos_a == A.pos_b) ? 1 : 0;
  }
}

}  // namespace sis
