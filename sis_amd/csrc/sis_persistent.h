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
  int *huge;                // a conflict component above the bucket capacity this step
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
__device__ __forceinline__ int bp_eval_one(const Params &P, const Ptrs &D, int gb, int mc,
                                           const double4 *xr = nullptr, const int *mpl = nullptr,
                                           int *ntc = nullptr) {
  const int *mp = mpl ? mpl : D.bp_mp + (size_t)gb * 8;   // mpl: the same record staged in LDS
  const int r = mp[7];
  const int N = P.N;
  D.bp_status[gb] = 0;
  D.ene_bp[gb] = 0.0;
  if (!mc && P.nstep_bp_mc > 0 && !D.bp_status_mc[gb]) return 0;
  double tK = D.tK[r];
  double dG = 1.0;   // models 1/3: U0 exp(-u) (src/force_bp.F90:203)
  if (P.bp_model == 5) {
    dG = D.bp_coef[2 * (size_t)gb] - tK * D.bp_coef[2 * (size_t)gb + 1];
    if (dG >= 0.0) return 0;
  }
  const double *box = D.box + 3 * r;
  const double4 *x = xr ? xr : D.pos + (size_t)r * N;
  double4 p1 = x[mp[0] - 1], p2 = x[mp[1] - 1];
  V3<double> v12 = pvec(P, box, p1, p2), v13 = pvec(P, box, p1, x[mp[2] - 1]),
             v15 = pvec(P, box, p1, x[mp[4] - 1]), v42 = pvec(P, box, x[mp[3] - 1], p2),
             v62 = pvec(P, box, x[mp[5] - 1], p2);
  const BpPara &bp = D.bpp[mp[6]];
  T u;
  V3<T> fb[6];
  if (!bp_geom<T>(bp, v12, v13, v15, v42, v62, u, fb)) return 0;
  double e = bp.U0 * dG * exp(-(double)u);
  if (!(e <= P.bp_cutoff_energy)) return 0;   // also rejects a NaN energy
  D.bp_status[gb] = 1;
  D.ene_bp[gb] = e;
  double *fo = D.for_bp + (size_t)gb * 18;
  for (int q = 0; q < 6; q++) {
    fo[3 * q] = e * (double)fb[q].x;
    fo[3 * q + 1] = e * (double)fb[q].y;
    fo[3 * q + 2] = e * (double)fb[q].z;
  }
  int *cnt = ntc ? ntc : D.nt_cnt + (size_t)r * N;   // ntc: this replica's counts in LDS
  atomicAdd(&cnt[mp[0] - 1], 1);
  atomicAdd(&cnt[mp[1] - 1], 1);
  return 1;
}

template <typename T, int LPB, int B>
__global__ void __launch_bounds__(B) k_persistent(Params P, Ptrs D, PersArgs A) {
  __shared__ double s1[B / LPB], s2[B / LPB];
  __shared__ int s_exit, s_big;
  __shared__ int s_mcws[(B / 64) * MCW_LDS];   // wave MC scratch, one slice per wave
  static_assert((B / 64) * MCW_LDS >= MCB_LDS, "large-component MC needs the block's whole scratch");
  if (threadIdx.x == 0) s_big = 0;
  const int N = P.N;
  const unsigned G = gridDim.x;
  const int gtid = blockIdx.x * B + threadIdx.x, gthreads = G * B;
  unsigned long long target = 0;
  const double beta = mc_beta(P, D.tK[0], true);
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
    // bonded forces read only positions: computed here, added in phase D
    {
      const int grp = gtid / LPB, l = gtid % LPB, ngrp = gthreads / LPB;
      for (int a0 = 0; a0 < N; a0 += ngrp) {
        int a = a0 + grp;
        V3<double> F{0.0, 0.0, 0.0};
        if (a < N) F = bonded_part(P, D, 0, a, l, LPB);
        F = group_sum<LPB>(F);
        if (a < N && l == 0) D.fbond[a] = double4{F.x, F.y, F.z, 0.0};
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
      // ---- C: small components one lane each, larger ones one wavefront each
      {
        for (int root = gtid; root < N; root += gthreads) {
          if (!A.has_exc[root]) continue;
          const int n = A.seg_n[root];
          if (n > MCL_PAIRS) continue;
          int *sg = A.seg + (size_t)root * SEGCAP;
          if (!mc_component_lane(D, N, 0, 0, root, sg, n, beta, A.seed, step))
            mc_component_serial(mc_arrays(D), P.N, 0, 0, root, sg, n, beta, A.seed, step);
        }
        const int wave = gtid / 64, nwaves = gthreads / 64;
        for (int root = wave; root < N; root += nwaves) {
          if (!A.has_exc[root]) continue;
          int n = A.seg_n[root];
          if (n <= MCL_PAIRS) continue;
          if (n > SEGCAP) {   // resolved after the phase by mc_component_huge
            if ((threadIdx.x & 63) == 0) *A.huge = 1;
            continue;
          }
          if (n > 64 * MCW_ITEMS) { s_big = 1; continue; }
          int *sg = A.seg + (size_t)root * SEGCAP;
          if (!mc_component_wave_any(P, D, 0, 0, root, sg, n, beta, A.seed, step, s_mcws + (threadIdx.x / 64) * MCW_LDS)) {
            if ((threadIdx.x & 63) == 0) mc_component_serial(mc_arrays(D), P.N, 0, 0, root, sg, n, beta, A.seed, step);
          }
        }
        // components above one wave's scratch: wave 0 of the block takes
        // those its waves met, with the whole block's scratch
        __syncthreads();
        if (s_big) {
          if (threadIdx.x < 64) {
            const int L = threadIdx.x;
            for (int lw = 0; lw < B / 64; lw++) {
              const int w0 = blockIdx.x * (B / 64) + lw;
              for (int k0 = 0; w0 + k0 * nwaves < N; k0 += 64) {
                const int root = w0 + (k0 + L) * nwaves;
                bool big = false;
                if (root < N) {
                  const int n = A.seg_n[root];
                  big = A.has_exc[root] && n > 64 * MCW_ITEMS && n <= SEGCAP;
                }
                for (unsigned long long m = __ballot(big); m; m &= m - 1) {
                  const int rt = bcast(root, __ffsll((long long)m) - 1);
                  const int n = A.seg_n[rt];
                  int *sg = A.seg + (size_t)rt * SEGCAP;
                  if (!mc_component_wave<MCB_ITEMS, MCB_NODES>(P, D, 0, 0, rt, sg, n, beta, A.seed, step, s_mcws))
                    if (L == 0) mc_component_serial(mc_arrays(D), P.N, 0, 0, rt, sg, n, beta, A.seed, step);
                }
              }
            }
          }
          __syncthreads();
          if (threadIdx.x == 0) s_big = 0;
        }
      }
      grid_barrier(A.bar, target, G);
      // components above the bucket capacity: wave 0 of block 0, then a
      // barrier (every block reads the same flag: it is written before the
      // barrier above and cleared after the one below)
      if (__hip_atomic_load(A.huge, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        if (blockIdx.x == 0 && threadIdx.x < 64) {
          const int L = threadIdx.x;
          for (int u0 = 0; u0 < N; u0 += 64) {
            const int root = u0 + L;
            const bool h = root < N && A.has_exc[root] && A.seg_n[root] > SEGCAP;
            for (unsigned long long m = __ballot(h); m; m &= m - 1) {
              const int rt = bcast(root, __ffsll((long long)m) - 1);
              if (!mc_component_huge(mc_arrays(D), N, 0, 0, rt, D.comp, 0, A.nbp_total, beta, A.seed, step, A.seg,
                                     (size_t)N * SEGCAP, s_mcws))
                if (L == 0) D.flags[5] = 1;
            }
          }
        }
        grid_barrier(A.bar, target, G);
        if (blockIdx.x == 0 && threadIdx.x == 0) *A.huge = 0;
      }
    }
    // ---- D: forces and the update; reset per-step MC state.  The MC
    // result is what later non-MC steps evaluate (src/job_md.F90:388-395).
    if (mc)
      for (int gb = gtid; gb < A.nbp_total; gb += gthreads) D.bp_status_mc[gb] = D.bp_status[gb];
    double m1 = 0.0, m2 = 0.0;
    {
      const int grp = gtid / LPB, l = gtid % LPB, ngrp = gthreads / LPB;
      for (int a0 = 0; a0 < N; a0 += ngrp) {
        int a = a0 + grp;
        V3<double> F{0.0, 0.0, 0.0};
        if (a < N) F = pair_part<T>(P, D, 0, a, l, LPB);
        F = group_sum<LPB>(F);
        if (a < N && l == 0) {
          const double4 fb = D.fbond[a];
          F.x += fb.x; F.y += fb.y; F.z += fb.z;
          D.nt_cnt[a] = -P.max_bp_per_nt;
          D.comp[a] = a;
          A.seg_n[a] = 0;
          A.has_exc[a] = 0;
          double d2 = gjf_update(P, D, 0, 0, a, F, 0, A.seed, step);
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

namespace sis {

// ---------------------------------------------------------------------
// Batched form (mode 3) and the job pool.  One workgroup advances one
// system through whole job_md steps with workgroup barriers only,
// rebuilding its own lists when its skin test fires.  batched_body works
// on replica r of (P, D): k_batched runs replica r = blockIdx.x of one
// context; k_jobs runs one job per workgroup with r = 0 and (P, D)
// pointing at that job's own arrays.  Candidate ids are r * capb + k.

constexpr int EXC_LDS = 1024;   // listed roots with excess per system and step
constexpr int MCW_LIST = 64;    // wavefront-MC components scheduled largest first per step
// Positions of systems up to this many beads are read from an LDS copy in
// the one-workgroup forms (32 KB of dynamic LDS beside about 31 KB static).
constexpr int BATCH_X_LDS = 1024;
// Per-lane partial pair forces (bead count x lanes per bead) kept in LDS
// beside the positions, 16 KB at most.
constexpr int BATCH_FP_LDS = 512;
__host__ __device__ inline int lanes_per_bead(int N) { return N <= 32 ? 8 : (N <= 64 ? 4 : (N <= 128 ? 2 : 1)); }

// Per-system launch arguments.
struct SysIO {
  int todo;          // steps to run in this launch
  int mc_first;      // limit MC on the first step (lists rebuilt by the host)
  int capb;          // candidate slots of this system
  int lpb;           // lanes per bead: 1, 2, 4 or 8
  int rk;            // key of the system's random streams
  int64_t stepno;    // istep before the launch
  uint64_t seed;
  int *seg_n, *seg, *has_exc;   // MC scratch (replica-r indexing)
  double4 *pos_b;               // second position buffer (replica-r indexing)
  int *done, *nbp, *rebuild, *nreb;  // this system's slots
  int xlds;                     // beads whose positions fit the launch's dynamic LDS
  int fplds;                    // per-lane partial forces that fit after them
  int nblds;                    // neighbour-list entries that fit after those (0: lists read from HBM)
  int stlds;                    // 1: the dynamics state lives in LDS for the launch (after the partial forces)
  int mplds;                    // ints of LDS for the candidate records after that (0: none)
  int force_only;               // evaluate the forces of the current positions into D.frc, no update
  // energy mode: one evaluation of energy_sumup at this system's D.tK (BP
  // terms and the limit MC at that temperature), DH coefficient e_coef and
  // Debye length e_lam, into energy_out[0..15]; nothing moves, the MC
  // state of the steps (bp_status_mc) is left alone
  int energy_mode;
  double e_coef, e_lam;
  double *energy_out;
  int64_t *stepno_out;          // istep after the launch (device-driven segments), or null
  int *mc_first_out;            // cleared after the launch, or null
};

// Launch arguments of k_batched (arrays over replicas).
struct BatchArgs {
  const int *todo;
  int64_t *stepno;
  int *mc_first;
  uint64_t seed;
  int *seg_n, *seg, *has_exc;
  int *done, *nbp, *rebuild, *nreb;
  double4 *pos_b;
  int capb, lpb;
  int xlds;   // beads whose positions fit the dynamic LDS of the launch
  int fplds;  // per-lane partial forces that fit after them
  int nblds;  // neighbour-list entries that fit after those (0: lists read from HBM)
  int stlds;  // 1: velocities, noise terms, displacements, forces, next positions and bonded
              // forces in LDS for the launch, written back at its end
  int mplds;  // ints of LDS for the replica's candidate records (8 each) after that, or 0
  int force_only;   // one force evaluation (limit MC if mc_first) into D.frc, no update
  // energy mode (one energy_sumup per replica, see SysIO): DH coefficient
  // and Debye length per replica (null: the replica's own), output
  // energy[r * 2 + energy_slot], and replicas to skip (their output 0)
  int energy_mode, energy_slot;
  const double *e_coef, *e_lam;
  const int *e_use;
  double *energy;
  int write_back;   // write istep and clear mc_first on the device after the launch
};

// Base-pair candidate arrays of one layout.
struct BpArrays {
  int *mp;
  double *coef;
  uint8_t *status, *status_mc;
  double *ene, *for_bp;
};

// Candidates move between the packed layout (replicas concatenated in the
// reference's list order; replica r at rep_off[r]..rep_off[r+1]) and the
// batched one (replica r at r * capb + k).  dir = 1: packed -> batched.
__global__ void k_bp_relayout(BpArrays S, BpArrays T, const int *rep_off, int capb, int dir) {
  const int r = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = rep_off[r + 1] - rep_off[r];
  if (k >= n) return;
  const size_t pk = (size_t)rep_off[r] + k, bt = (size_t)r * capb + k;
  const size_t src = dir ? pk : bt, dst = dir ? bt : pk;
  for (int q = 0; q < 8; q++) T.mp[dst * 8 + q] = S.mp[src * 8 + q];
  T.coef[2 * dst] = S.coef[2 * src];
  T.coef[2 * dst + 1] = S.coef[2 * src + 1];
  T.status[dst] = S.status[src];
  T.status_mc[dst] = S.status_mc[src];
  T.ene[dst] = S.ene[src];
  for (int q = 0; q < 18; q++) T.for_bp[dst * 18 + q] = S.for_bp[src * 18 + q];
}

// Incidence entries (candidate * 8 + role) renumbered in place to the
// other layout; every entry of a bead of replica r names a candidate of r.
__global__ void k_inc_remap(Params P, Ptrs D, const int *rep_off, int capb, int dir) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (a >= P.N) return;
  const int N = P.N;
  const size_t g = (size_t)r * N + a;
  const int n = min(D.n_inc[g], D.maxinc);
  int *inc = D.inc + (size_t)r * D.maxinc * N + a;
  for (int k = 0; k < n; k++) {
    const int e = inc[(size_t)k * N], gb = e >> 3, role = e & 7;
    const int ngb = dir ? r * capb + (gb - rep_off[r]) : rep_off[r] + (gb - r * capb);
    inc[(size_t)k * N] = ngb * 8 + role;
  }
}

__device__ __forceinline__ V3<double> group_sum_rt(V3<double> F, int lpb) {
  for (int off = lpb >> 1; off > 0; off >>= 1) {
    F.x += __shfl_xor(F.x, off, lpb);
    F.y += __shfl_xor(F.y, off, lpb);
    F.z += __shfl_xor(F.z, off, lpb);
  }
  return F;
}

// neighbor_list for one system inside its workgroup: pbc_wrap, all-pairs
// search over LDS position tiles with the criteria of
// src/neighbor_list.F90:123-223, candidates in owner-bead order, their
// incidences sorted per bead.  Sets *s_flag when a capacity is exceeded.
template <int B>
__device__ __forceinline__ void batched_rebuild(const Params &P, const Ptrs &D, const SysIO &J, int r, int *s_scan,
                                double4 *s_tile, int *s_flag) {
  const int N = P.N;
  const int tid = threadIdx.x;
  const size_t base = (size_t)r * N;
  // pbc_wrap (src/pbc.F90:196-226)
  if (P.pbc) {
    const double *box = D.box + 3 * r;
    for (int a = tid; a < N; a += B) {
      double4 &p = D.pos[base + a];
      double v[3] = {p.x, p.y, p.z};
      for (int k = 0; k < 3; k++) {
        double L = box[k];
        if (v[k] > L) v[k] = v[k] - L * ((double)(int)((v[k] - L) / L) + 1);
        else if (v[k] < 0.0) v[k] = v[k] + L * ((double)(int)((0.0 - v[k]) / L) + 1);
      }
      p.x = v[0]; p.y = v[1]; p.z = v[2];
    }
  }
  // the electrostatic cutoff of the new lists (src/neighbor_list.F90:80-82)
  const double ecut = rebuild_ele_cut(P, D, r);
  __syncthreads();
  if (tid == 0) D.ele_cut[r] = ecut;
  const double ec = ecut + P.nl_margin;
  const double ele_cut2 = P.ele_on ? ec * ec : 0.0;
  for (int a0 = 0; a0 < N; a0 += B) {
    const int a = a0 + tid;
    const bool act = a < N;
    SearchOut o{0, 0, 0};
    const double4 pa = act ? D.pos[base + a] : double4{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < N; t0 += B) {
      __syncthreads();
      if (t0 + tid < N) s_tile[tid] = D.pos[base + t0 + tid];
      __syncthreads();
      if (act) {
        const int nt = min(B, N - t0);
        for (int k = 0; k < nt; k++) {
          const int b = t0 + k;
          if (b != a) pair_test(P, D, r, a, b, pa, s_tile[k], ele_cut2, o);
        }
      }
    }
    if (act) {
      if (o.ne > D.maxnb || o.nw > D.maxw || o.nb > D.maxbpb) *s_flag = 1;
      search_finish(P, D, r, a, o);
      D.n_inc[base + a] = 0;
    }
  }
  __syncthreads();
  // candidate offsets: exclusive scan of the per-bead counts
  int running = 0;
  for (int a0 = 0; a0 < N; a0 += B) {
    const int a = a0 + tid;
    const int v = a < N ? D.n_bp_tmp[base + a] : 0;
    s_scan[tid] = v;
    __syncthreads();
    for (int off = 1; off < B; off <<= 1) {
      const int t = tid >= off ? s_scan[tid - off] : 0;
      __syncthreads();
      s_scan[tid] += t;
      __syncthreads();
    }
    if (a < N) D.bp_off[base + a] = running + s_scan[tid] - v;
    running += s_scan[B - 1];
    __syncthreads();
  }
  const int nbr = min(running, J.capb);
  if (tid == 0) {
    if (running > J.capb) *s_flag = 1;
    *J.nbp = nbr;
  }
  for (int a = tid; a < N; a += B) {
    const size_t g = base + a;
    const Topo ti = D.topo[a];
    const int nb = D.n_bp_tmp[g], off = D.bp_off[g];
    for (int k = 0; k < nb; k++) {
      const int lk = off + k;
      if (lk >= J.capb) break;
      const int gb = r * J.capb + lk;
      const int b = D.bp_tmp[g * D.maxbpb + k];
      const Topo tj = D.topo[b];
      const int i = ti.loc, j = tj.loc, imp = a + 1, jmp = b + 1;
      int *dst = D.bp_mp + (size_t)gb * 8;
      dst[0] = imp;
      dst[1] = jmp;
      dst[2] = (i == 1) ? imp + ti.len - 1 : imp - 1;
      dst[3] = (j == 1) ? jmp + tj.len - 1 : jmp - 1;
      dst[4] = (i == ti.len) ? imp + 1 - ti.len : imp + 1;
      dst[5] = (j == tj.len) ? jmp + 1 - tj.len : jmp + 1;
      int bt;
      const int h = bp_eligible(P, D, a, b, &bt);
      dst[6] = bt;
      dst[7] = r;
      D.bp_coef[2 * (size_t)gb] = P.bp_model == 5 ? D.bp3_dH[h] : 0.0;
      D.bp_coef[2 * (size_t)gb + 1] = P.bp_model == 5 ? D.bp3_dS[h] : 0.0;
    }
  }
  __syncthreads();
  for (int k = tid; k < nbr; k += B) {
    const int gb = r * J.capb + k;
    const int *mp = D.bp_mp + (size_t)gb * 8;
    for (int q = 0; q < 6; q++) {
      const size_t gq = base + (mp[q] - 1);
      const int slot = atomicAdd(&D.n_inc[gq], 1);
      if (slot < D.maxinc) D.inc[((size_t)r * D.maxinc + slot) * N + (mp[q] - 1)] = gb * 8 + q;
      else *s_flag = 1;
    }
  }
  __syncthreads();
  for (int a = tid; a < N; a += B) {
    const size_t g = base + a;
    const int n = min(D.n_inc[g], D.maxinc);
    const size_t ib = (size_t)r * D.maxinc * N + a, st = N;
    for (int x = 1; x < n; x++) {
      const int v = D.inc[ib + x * st];
      int y = x - 1;
      while (y >= 0 && D.inc[ib + y * st] > v) { D.inc[ib + (y + 1) * st] = D.inc[ib + y * st]; y--; }
      D.inc[ib + (y + 1) * st] = v;
    }
    D.move[g] = double4{0.0, 0.0, 0.0, 0.0};
    D.nt_cnt[g] = -P.max_bp_per_nt;
    D.comp[g] = a;
    J.seg_n[g] = 0;
    J.has_exc[g] = 0;
  }
  __syncthreads();
}

// Dynamic LDS of one batched launch: positions, per-lane partial forces,
// (stlds) six arrays of N double4 of dynamics state and four of N ints of
// limit-MC scratch, mplds ints of candidate records, then (nblds > 0)
// per-bead list offsets and counts and the list entries.
constexpr int BATCH_ST_ARRAYS = 6, BATCH_ST_INTS = 4;
__host__ __device__ inline size_t batched_lds_bytes(int xlds, int fplds, int nblds, int N, int stlds = 0,
                                                    int mplds = 0) {
  return (size_t)(xlds + fplds + (stlds ? BATCH_ST_ARRAYS * N : 0)) * 16 * 2 +
         (size_t)((stlds ? BATCH_ST_INTS * N : 0) + mplds) * sizeof(int) +
         (nblds > 0 ? (size_t)(2 * N + nblds) * sizeof(int) : 0);
}

// pair_nb with the replica's lists staged in LDS: bead a's WCA entries at
// s_nb[off ..), its Debye-Hueckel entries right after them (cnt = wca
// count << 16 | DH count).  Same pairs in the same order as pair_nb.
template <typename T>
__device__ __forceinline__ V3<double> pair_nb_lds(const Params &P, const Ptrs &D, int r, int a, int l, int LPB,
                                                  const double4 *x, const int *s_nb, int off, int cnt) {
  const double *box = D.box + 3 * r;
  const double4 pa = x[a];
  const int nw = cnt >> 16, ne = P.ele_on ? (cnt & 0xffff) : 0;
  const T coef = (T)D.ele_coef[r], rc2 = (T)(D.ele_cut[r] * D.ele_cut[r]), rl = (T)(1.0 / D.lambdaD[r]);
  V3<double> F{0.0, 0.0, 0.0};
  for (int k = l; k < nw; k += LPB) F = F + cvt<double>(wca_force<T>(P, cvt<T>(pvec(P, box, pa, x[s_nb[off + k]]))));
  const int *e = s_nb + off + nw;
  for (int k = l; k < ne; k += LPB) F = F + cvt<double>(dh_force<T>(coef, rc2, rl, cvt<T>(pvec(P, box, pa, x[e[k]]))));
  return F;
}

// Whole job_md steps of system r (src/job_md.F90:374-460) for J.todo steps.
template <typename T, int B>
__device__ __forceinline__ void batched_body(const Params &P, const Ptrs &D, const SysIO &J, const int r) {
  __shared__ double s1[B], s2[B];
  __shared__ int s_scan[B];
  __shared__ double4 s_tile[B];
  __shared__ int s_flag, s_exit;
  __shared__ int s_exc[EXC_LDS];   // component roots with excess (MC steps)
  __shared__ int s_nexc, s_big, s_nw, s_next, s_huge;
  __shared__ int s_wl[MCW_LIST];   // wavefront-MC components of this step: size << 21 | root
  __shared__ int s_mcws[(B / 64) * MCW_LDS];   // wave MC scratch, one slice per wave
  static_assert((B / 64) * MCW_LDS >= MCB_LDS, "large-component MC needs the block's whole scratch");
  const int N = P.N;
  const int tid = threadIdx.x;
  const size_t base = (size_t)r * N;
  const int lpb = J.lpb, grp = tid / lpb, l = tid % lpb, ngrp = B / lpb;
  const double beta = mc_beta(P, D.tK[r], true);
  // the step reads positions from an LDS copy when they fit (refreshed
  // after every update and every rebuild, which may wrap them)
  extern __shared__ double4 s_x[];
  const bool xl = N <= J.xlds;
  const double4 *xr = xl ? s_x : D.pos + base;
  // per-lane partial pair forces (WCA + Debye-Hueckel) follow the positions
  // in the dynamic LDS when they fit; the work loop of phase C fills them
  double4 *s_fp = s_x + J.xlds;
  const bool ovl = xl && N * lpb <= J.fplds;
  auto stage_x = [&]() {
    if (xl)
      for (int a = tid; a < N; a += B) s_x[a] = D.pos[base + a];
    __syncthreads();
  };
  // the dynamics state of the replica in LDS for the whole launch when it
  // fits (MD launches only): velocities, noise terms, accumulated
  // displacements, forces, the next positions and the bonded forces.  DS is
  // D with those arrays (and the current positions, s_x) pointing into LDS,
  // indexed by bead alone (replica slot rs = 0, offset sb = 0; the one
  // per-replica global array the update reads, md_cr, is shifted to this
  // replica instead).  LDS pointers are never offset below their base: a
  // flat pointer into LDS keeps its offset in the low 32 bits.  The global
  // arrays are read at the start and written back before a rebuild and at
  // the end.
  const bool st = xl && J.stlds && !J.energy_mode && !J.force_only;
  double4 *s_st = s_x + J.xlds + J.fplds;
  // with the state, the per-step limit-MC scratch of the replica (component
  // parents, bucket sizes, excess flags, per-nucleotide pair counts; N ints
  // each) lives in LDS too.  The MC routines index the counts by replica:
  // they get DM, D with the counts pointing at the LDS row, and replica
  // slot rm = 0.
  int *s_sti = reinterpret_cast<int *>(s_st + BATCH_ST_ARRAYS * N);
  int *comp = st ? s_sti : D.comp + base;
  int *segn = st ? s_sti + N : J.seg_n + base;
  int *hexc = st ? s_sti + 2 * N : J.has_exc + base;
  int *ntc = st ? s_sti + 3 * N : D.nt_cnt + base;
  Ptrs DM = D;
  int rm = r;
  if (st) {
    DM.nt_cnt = ntc;
    rm = 0;
  }
  Ptrs DS = D;
  size_t sb = base;
  int rs = r;
  if (st) {
    DS.vel = s_st;
    DS.acc = s_st + N;
    DS.move = s_st + 2 * N;
    DS.frc = s_st + 3 * N;
    DS.pos_out = s_st + 4 * N;
    DS.fbond = s_st + 5 * N;
    DS.pos = s_x;
    DS.md_cr = D.md_cr + base;
    sb = 0;
    rs = 0;
  }
  auto state_in = [&]() {   // global -> LDS (positions are staged by stage_x)
    if (st)
      for (int a = tid; a < N; a += B) {
        s_st[a] = D.vel[base + a];
        s_st[N + a] = D.acc[base + a];
        s_st[2 * N + a] = D.move[base + a];
      }
  };
  auto state_out = [&]() {   // LDS -> global: positions and displacements (for a rebuild)
    if (st)
      for (int a = tid; a < N; a += B) {
        D.pos[base + a] = s_x[a];
        D.move[base + a] = s_st[2 * N + a];
      }
  };
  // the replica's base-pair candidate records (bp_mp, 8 ints each) in LDS
  // after the state when they fit (s_mpok; they change only at rebuilds)
  __shared__ int s_mpok;
  int *s_mp = J.stlds ? s_sti + BATCH_ST_INTS * N : reinterpret_cast<int *>(s_st);
  // the replica's WCA and Debye-Hueckel lists in LDS after those, when they
  // fit (read by every step until the next rebuild): bead a's entries from
  // s_noff[a], counts packed in s_ncnt[a]
  __shared__ int s_nbok;
  int *s_ncnt = s_mp + J.mplds;
  int *s_noff = s_ncnt + N, *s_nb = s_noff + N;
  auto stage_lists = [&]() {
    int running = 0;
    if (xl && J.nblds > 0) {
      for (int a0 = 0; a0 < N; a0 += B) {
        const int a = a0 + tid;
        int cw = 0, ce = 0;
        if (a < N) {
          cw = min(D.n_wca[base + a], D.maxw);
          ce = P.ele_on ? min(D.n_ele[base + a], D.maxnb) : 0;
        }
        const int v = cw + ce;
        s_scan[tid] = v;
        __syncthreads();
        for (int off = 1; off < B; off <<= 1) {
          const int t = tid >= off ? s_scan[tid - off] : 0;
          __syncthreads();
          s_scan[tid] += t;
          __syncthreads();
        }
        if (a < N) {
          s_noff[a] = running + s_scan[tid] - v;
          s_ncnt[a] = (cw << 16) | ce;
        }
        running += s_scan[B - 1];
        __syncthreads();
      }
    }
    const bool ok = xl && J.nblds > 0 && running <= J.nblds;
    if (ok) {
      // slot k of every bead in turn: coalesced reads of the k-major lists
      const int *nbw = D.nb_wca + (size_t)r * D.maxw * N, *nbe = D.nb_ele + (size_t)r * D.maxnb * N;
      const int kmax = max(D.maxw, P.ele_on ? D.maxnb : 0);
      for (int i = tid; i < kmax * N; i += B) {
        const int k = i / N, a = i - k * N;
        const int c = s_ncnt[a], cw = c >> 16, ce = c & 0xffff;
        if (k < cw) s_nb[s_noff[a] + k] = nbw[(size_t)k * N + a];
        if (k < ce) s_nb[s_noff[a] + cw + k] = nbe[(size_t)k * N + a];
      }
    }
    const int nbc = *J.nbp;
    const bool mok = J.mplds > 0 && nbc * 8 <= J.mplds;
    if (mok) {
      const int *src = D.bp_mp + (size_t)r * J.capb * 8;
      for (int i = tid; i < nbc * 8; i += B) s_mp[i] = src[i];
    }
    if (tid == 0) {
      s_nbok = ok ? 1 : 0;
      s_mpok = mok ? 1 : 0;
    }
    __syncthreads();
  };
  if (tid == 0) s_flag = 0;
  // per-step limit-MC state starts reset whatever ran before this launch
  for (int a = tid; a < N; a += B) {
    ntc[a] = -P.max_bp_per_nt;
    comp[a] = a;
    segn[a] = 0;
    hexc[a] = 0;
  }
  __syncthreads();
  // the first step after a rebuild does the limit MC (src/job_md.F90:388-395)
  int mc_next = J.mc_first;
  int nreb = 0;
  if (J.todo > 0 && *J.rebuild && !J.force_only && !J.energy_mode) {
    batched_rebuild<B>(P, D, J, r, s_scan, s_tile, &s_flag);
    if (tid == 0) *J.rebuild = 0;
    mc_next = 1;
    nreb++;
  }
  stage_x();
  state_in();
  stage_lists();
  int64_t completed = 0;
  int overflow = s_flag;
  const int64_t nsteps = J.todo, step0 = J.stepno;
  for (int64_t s = 0; s < nsteps && !overflow; s++) {
    // the energy evaluation keys its MC draws by the current step, as the
    // host's energy_sumup after that step does
    const int64_t step = J.energy_mode ? step0 : step0 + s + 1;
    const int mc = J.energy_mode ? P.nstep_bp_mc > 0
                   : J.force_only ? (P.nstep_bp_mc > 0 && J.mc_first)
                                  : P.nstep_bp_mc > 0 && ((step % P.nstep_bp_mc) == 0 || mc_next);
    mc_next = 0;
    const int nbr = *J.nbp;
    // ---- A: base pairs (+ union of formed pairs on MC steps); bonded
    // forces here unless the work loop below takes them
    for (int k = tid; k < nbr; k += B) {
      const int gb = r * J.capb + k;
      const int *mpl = s_mpok ? s_mp + 8 * k : nullptr;
      if (bp_eval_one<T>(P, D, gb, mc, xr, mpl, st ? ntc : nullptr) && mc) {
        const int *mp = mpl ? mpl : D.bp_mp + (size_t)gb * 8;
        int a = mp[0] - 1, b = mp[1] - 1;
        while (true) {
          a = uf_find(comp, a);
          b = uf_find(comp, b);
          if (a == b) break;
          if (a > b) { int t = a; a = b; b = t; }
          if (atomicCAS(&comp[b], b, a) == b) break;
        }
      }
    }
    if (!ovl)
      for (int a0 = 0; a0 < N; a0 += ngrp) {
        const int a = a0 + grp;
        V3<double> F{0.0, 0.0, 0.0};
        if (a < N) F = bonded_part(P, D, r, a, l, lpb, xr);
        F = group_sum_rt(F, lpb);
        if (a < N && l == 0) DS.fbond[sb + a] = double4{F.x, F.y, F.z, 0.0};
      }
    if (tid == 0) { s_nexc = 0; s_big = 0; s_nw = 0; s_next = 0; s_huge = 0; }
    __syncthreads();
    int nw = 0, nexc = 0;
    bool listed = true;
    if (mc) {
      // ---- B: formed pairs bucketed by component root; roots with
      // excess listed once
      for (int k = tid; k < nbr; k += B) {
        const int gb = r * J.capb + k;
        if (!D.bp_status[gb]) continue;
        const int *mp = s_mpok ? s_mp + 8 * k : D.bp_mp + (size_t)gb * 8;
        const int a = mp[0] - 1, b = mp[1] - 1;
        const int root = uf_find(comp, a);
        const size_t gr = base + root;
        const int slot = atomicAdd(&segn[root], 1);
        if (slot < SEGCAP) J.seg[gr * SEGCAP + slot] = gb;
        if ((ntc[a] > 0 || ntc[b] > 0) && atomicExch(&hexc[root], 1) == 0) {
          const int i = atomicAdd(&s_nexc, 1);
          if (i < EXC_LDS) s_exc[i] = root;
        }
      }
      __syncthreads();
      // ---- C: components with excess classified by size (components are
      // disjoint and keyed by their root, so the order does not matter):
      // the wavefront ones listed with their size, the large ones flagged;
      // the small ones (one lane each) are work-list items below
      const int wave = tid / 64;
      nexc = s_nexc;
      listed = nexc <= EXC_LDS;
      for (int it = tid; it < (listed ? nexc : N); it += B) {
        const int root = listed ? s_exc[it] : it;
        if (!hexc[root]) continue;
        const int n = segn[root];
        if (n > SEGCAP) { s_huge = 1; continue; }
        if (n > 64 * MCW_ITEMS) { s_big = 1; continue; }
        if (n > MCL_PAIRS) {
          const int q = atomicAdd(&s_nw, 1);
          if (q < MCW_LIST) s_wl[q] = (n << 21) | root;   // root < 2^21 (N <= 2048 here)
        }
      }
      __syncthreads();
      nw = s_nw;
      if (nw <= MCW_LIST) {
        // largest first (rank by size, then root: keys are distinct)
        if (wave == 0) {
          const int L = tid & 63;
          const int key = L < nw ? s_wl[L] : 0;
          int rank = 0;
          for (int j = 0; j < nw; j++) rank += s_wl[j] > key;
          wave_lds_fence();
          if (L < nw) s_wl[rank] = key;
        }
        __syncthreads();
      }
    }
    // ---- C: one work list for the waves, handed out one item at a time:
    // the wavefront-MC components (largest first), then the small
    // components, one lane each, 64 listed roots an item, then, when the
    // partial forces fit in LDS, the bonded and WCA/Debye-Hueckel forces of
    // groups of beads (they do not depend on the MC).  The longest wave is
    // then close to the largest component, with the rest in its shadow.
    {
      const int wave = tid / 64, L = tid & 63;
      const int nmc = (mc && nw <= MCW_LIST) ? nw : 0;
      const int nroot = listed ? nexc : N;
      const int nlane = mc ? (nroot + 63) / 64 : 0;
      const int bpw = 64 / lpb, nchunk = ovl ? (N + bpw - 1) / bpw : 0;
      for (;;) {
        int k = 0;
        if (L == 0) k = atomicAdd(&s_next, 1);
        k = bcast(k, 0);
        if (k >= nmc + nlane + nchunk) break;
        if (k < nmc) {
          const int key = s_wl[k];
          const int root = key & 0x1fffff, n = key >> 21;
          int *sg = J.seg + (base + root) * SEGCAP;
          if (!mc_component_wave_any(P, DM, rm, J.rk, root, sg, n, beta, J.seed, step, s_mcws + wave * MCW_LDS))
            if (L == 0) mc_component_serial(mc_arrays(DM), P.N, rm, J.rk, root, sg, n, beta, J.seed, step);
          continue;
        }
        if (k < nmc + nlane) {
          const int it = (k - nmc) * 64 + L;
          if (it < nroot) {
            const int root = listed ? s_exc[it] : it;
            const size_t gr = base + root;
            if (hexc[root]) {
              const int n = segn[root];
              if (n <= MCL_PAIRS) {
                int *sg = J.seg + gr * SEGCAP;
                if (!mc_component_lane(DM, N, rm, J.rk, root, sg, n, beta, J.seed, step))
                  mc_component_serial(mc_arrays(DM), P.N, rm, J.rk, root, sg, n, beta, J.seed, step);
              }
            }
          }
          continue;
        }
        const int a = (k - nmc - nlane) * bpw + L / lpb, lw = L % lpb;
        V3<double> F{0.0, 0.0, 0.0};
        if (a < N) {
          const V3<double> Fp = s_nbok ? pair_nb_lds<T>(P, D, r, a, lw, lpb, s_x, s_nb, s_noff[a], s_ncnt[a])
                                       : pair_nb<T>(P, D, r, a, lw, lpb, xr, V3<double>{0.0, 0.0, 0.0});
          s_fp[(size_t)a * lpb + lw] = double4{Fp.x, Fp.y, Fp.z, 0.0};
          F = bonded_part(P, D, r, a, lw, lpb, xr);
        }
        F = group_sum_rt(F, lpb);
        if (a < N && lw == 0) DS.fbond[sb + a] = double4{F.x, F.y, F.z, 0.0};
      }
    }
    __syncthreads();
    if (mc) {
      const int wave = tid / 64, nwaves = B / 64;
      if (nw > MCW_LIST) {
        // more than the list holds: static round robin over the roots
        for (int it = wave; it < (listed ? nexc : N); it += nwaves) {
          const int root = listed ? s_exc[it] : it;
          const size_t gr = base + root;
          if (!hexc[root]) continue;
          const int n = segn[root];
          if (n <= MCL_PAIRS || n > 64 * MCW_ITEMS) continue;
          int *sg = J.seg + gr * SEGCAP;
          if (!mc_component_wave_any(P, DM, rm, J.rk, root, sg, n, beta, J.seed, step, s_mcws + wave * MCW_LDS))
            if ((tid & 63) == 0) mc_component_serial(mc_arrays(DM), P.N, rm, J.rk, root, sg, n, beta, J.seed, step);
        }
        __syncthreads();
      }
      // components above one wave's scratch: wave 0 with the whole block's
      if (s_big) {
        if (wave == 0) {
          const int cnt = listed ? nexc : N;
          for (int i0 = 0; i0 < cnt; i0 += 64) {
            const int it = i0 + (tid & 63);
            int root = 0;
            bool big = false;
            if (it < cnt) {
              root = listed ? s_exc[it] : it;
              const int n = segn[root];
              big = hexc[root] && n > 64 * MCW_ITEMS && n <= SEGCAP;
            }
            for (unsigned long long m = __ballot(big); m; m &= m - 1) {
              const int rt = bcast(root, __ffsll((long long)m) - 1);
              const size_t gr = base + rt;
              const int n = segn[rt];
              int *sg = J.seg + gr * SEGCAP;
              if (!mc_component_wave<MCB_ITEMS, MCB_NODES>(P, DM, rm, J.rk, rt, sg, n, beta, J.seed, step, s_mcws))
                if ((tid & 63) == 0) mc_component_serial(mc_arrays(DM), P.N, rm, J.rk, rt, sg, n, beta, J.seed, step);
            }
          }
        }
        __syncthreads();
      }
      // components above the bucket capacity: wave 0, global scratch (the
      // replica's bucket region, free once the passes above are done)
      if (s_huge) {
        if (wave == 0) {
          const int cnt = listed ? nexc : N;
          for (int i0 = 0; i0 < cnt; i0 += 64) {
            const int it = i0 + (tid & 63);
            int root = 0;
            bool huge = false;
            if (it < cnt) {
              root = listed ? s_exc[it] : it;
              huge = hexc[root] && segn[root] > SEGCAP;
            }
            for (unsigned long long m = __ballot(huge); m; m &= m - 1) {
              const int rt = bcast(root, __ffsll((long long)m) - 1);
              if (!mc_component_huge(mc_arrays(DM), N, rm, J.rk, rt, comp, r * J.capb, r * J.capb + nbr, beta, J.seed,
                                     step, J.seg + base * SEGCAP, (size_t)N * SEGCAP, s_mcws))
                if ((tid & 63) == 0) D.flags[5] = 1;
            }
          }
        }
        __syncthreads();
      }
      // the MC result is what later non-MC steps evaluate
      if (!J.energy_mode)
        for (int k = tid; k < nbr; k += B) D.bp_status_mc[r * J.capb + k] = D.bp_status[r * J.capb + k];
    }
    if (J.energy_mode) {
      // energy_sumup: per-bead terms summed as k_energy and k_energy_reduce
      // do (blocks of 256 beads, tree sum, blocks in order), so the result
      // equals the per-phase evaluation's
      // (128 threads: each holds beads t and t + 128, the tree's first level)
      static_assert(B == 128 || B >= 256, "energy sums need 128 or at least 256 threads");
      constexpr int H = B < 256 ? B : 256;
      __syncthreads();
      double acc[7] = {0, 0, 0, 0, 0, 0, 0};
      for (int a0 = 0; a0 < N; a0 += 256) {
        double e[7] = {0, 0, 0, 0, 0, 0, 0}, e2[7] = {0, 0, 0, 0, 0, 0, 0};
        if (tid < H && a0 + tid < N) bead_energy(P, D, r, a0 + tid, J.e_coef, 1.0 / J.e_lam, mc, xr, e);
        if (B < 256 && tid < H && a0 + tid + 128 < N)
          bead_energy(P, D, r, a0 + tid + 128, J.e_coef, 1.0 / J.e_lam, mc, xr, e2);
        for (int k = 1; k < 7; k++) {
          if (tid < H) s1[tid] = B < 256 ? e[k] + e2[k] : e[k];
          __syncthreads();
          for (int off = H / 2; off > 0; off >>= 1) {
            if (tid < off) s1[tid] += s1[tid + off];
            __syncthreads();
          }
          if (tid == 0) acc[k] += s1[0];
          __syncthreads();
        }
      }
      if (tid == 0) {
        double t = 0.0;
        for (int k = 1; k <= 6; k++) t += acc[k];
        acc[0] = t;
        for (int k = 0; k < 7; k++) J.energy_out[k] = acc[k];
        for (int k = 7; k < 16; k++) J.energy_out[k] = 0.0;
      }
      break;
    }
    // ---- D: base-pair forces onto the partial pair forces (or the whole
    // pair sum when those did not fit in LDS), bonded, GJF-2GJ, per-step
    // MC state reset
    double m1 = 0.0, m2 = 0.0;
    for (int a0 = 0; a0 < N; a0 += ngrp) {
      const int a = a0 + grp;
      V3<double> F{0.0, 0.0, 0.0};
      if (a < N) {
        if (ovl) {
          const double4 fp = s_fp[(size_t)a * lpb + l];
          F = pair_bp(P, D, r, a, l, lpb, V3<double>{fp.x, fp.y, fp.z});
        } else if (s_nbok) {
          F = pair_bp(P, D, r, a, l, lpb, pair_nb_lds<T>(P, D, r, a, l, lpb, s_x, s_nb, s_noff[a], s_ncnt[a]));
        } else {
          F = pair_part<T>(P, D, r, a, l, lpb, xr);
        }
      }
      F = group_sum_rt(F, lpb);
      if (a < N && l == 0) {
        const size_t g = base + a;
        const double4 fb = DS.fbond[sb + a];
        F.x += fb.x; F.y += fb.y; F.z += fb.z;
        ntc[a] = -P.max_bp_per_nt;
        comp[a] = a;
        segn[a] = 0;
        hexc[a] = 0;
        if (J.force_only) {
          D.frc[g] = double4{F.x, F.y, F.z, 0.0};
        } else {
          const double d2 = gjf_update(P, DS, rs, J.rk, a, F, 0, J.seed, step);
          if (d2 > m1) { m2 = m1; m1 = d2; }
          else if (d2 > m2) m2 = d2;
        }
      }
    }
    if (J.force_only) {   // forces of the current positions only: no update, no skin test
      __syncthreads();
      break;
    }
    // the replica's two largest displacements: within each wave by lane
    // shuffles (lanes other than a group's first hold zeros), then over the
    // waves; the top two of a set do not depend on the order of the merges
    {
      double a1 = m1, a2 = m2;
      for (int off = 32; off > 0; off >>= 1) {
        const double b1 = __shfl_xor(a1, off), b2 = __shfl_xor(a2, off);
        const double n2 = a1 > b1 ? (a2 > b1 ? a2 : b1) : (a1 > b2 ? a1 : b2);
        a1 = a1 > b1 ? a1 : b1;
        a2 = n2;
      }
      if ((tid & 63) == 0) { s1[tid >> 6] = a1; s2[tid >> 6] = a2; }
    }
    __syncthreads();
    if (tid == 0) {
      double a1 = s1[0], a2 = s2[0];
      for (int w = 1; w < B / 64; w++) {
        const double b1 = s1[w], b2 = s2[w];
        const double n2 = a1 > b1 ? (a2 > b1 ? a2 : b1) : (a1 > b2 ? a1 : b2);
        a1 = a1 > b1 ? a1 : b1;
        a2 = n2;
      }
      s_exit = (sqrt(a1) + sqrt(a2) > P.nl_margin) ? 1 : 0;
    }
    // the updated positions become current (every reader of D.pos in
    // this step has passed the barriers above)
    for (int a = tid; a < N; a += B) {
      if (st) {
        s_x[a] = s_st[4 * N + a];
      } else {
        const double4 xn = D.pos_out[base + a];
        D.pos[base + a] = xn;
        if (xl) s_x[a] = xn;
      }
    }
    __syncthreads();
    completed = s + 1;
    if (s_exit) {
      if (s + 1 < nsteps) {
        // the rebuild works on the global positions and displacements
        state_out();
        __syncthreads();
        batched_rebuild<B>(P, D, J, r, s_scan, s_tile, &s_flag);
        overflow = s_flag;
        mc_next = 1;
        nreb++;
        stage_x();
        if (st)
          for (int a = tid; a < N; a += B) s_st[2 * N + a] = D.move[base + a];
        stage_lists();
      } else if (tid == 0) {
        *J.rebuild = 1;   // rebuild at the start of the next launch
      }
    }
  }
  // the state back to HBM (forces only once a step has produced them)
  if (st)
    for (int a = tid; a < N; a += B) {
      D.pos[base + a] = s_x[a];
      D.vel[base + a] = s_st[a];
      D.acc[base + a] = s_st[N + a];
      D.move[base + a] = s_st[2 * N + a];
      if (completed > 0) D.frc[base + a] = s_st[3 * N + a];
    }
  if (tid == 0 && !J.energy_mode && !J.force_only) {
    *J.done = (int)completed;
    *J.nreb = J.stepno_out ? *J.nreb + nreb : nreb;   // device-driven segments add up
    if (overflow) D.flags[0] = 1;
    if (J.stepno_out) *J.stepno_out = step0 + completed;
    if (J.mc_first_out) *J.mc_first_out = 0;
  }
}

template <typename T, int B>
__global__ void __launch_bounds__(B) k_batched(Params P, Ptrs D, BatchArgs A) {
  const int r = blockIdx.x;
  SysIO J;
  J.todo = A.todo[r];
  J.mc_first = A.mc_first[r];
  J.capb = A.capb;
  J.lpb = A.lpb;
  J.rk = r;
  J.stepno = A.stepno[r];
  J.seed = A.seed;
  J.seg_n = A.seg_n;
  J.seg = A.seg;
  J.has_exc = A.has_exc;
  J.pos_b = A.pos_b;
  J.xlds = A.xlds;
  J.fplds = A.fplds;
  J.nblds = A.nblds;
  J.stlds = A.stlds;
  J.mplds = A.mplds;
  J.force_only = A.force_only;
  J.done = A.done + r;
  J.nbp = A.nbp + r;
  J.rebuild = A.rebuild + r;
  J.nreb = A.nreb + r;
  J.energy_mode = A.energy_mode;
  J.e_coef = A.e_coef ? A.e_coef[r] : D.ele_coef[r];
  J.e_lam = A.e_lam ? A.e_lam[r] : D.lambdaD[r];
  J.energy_out = A.energy ? A.energy + ((size_t)r * 2 + A.energy_slot) * 16 : nullptr;
  J.stepno_out = A.write_back ? A.stepno + r : nullptr;
  J.mc_first_out = A.write_back ? A.mc_first + r : nullptr;
  if (A.energy_mode || A.write_back) {
    // a device-driven sequence: a capacity overflow earlier in it stops
    // every later launch (the host finishes the sequence on its own path)
    if (D.flags[0]) return;
    if (A.energy_mode && A.e_use && !A.e_use[r]) {
      if (threadIdx.x < 16) J.energy_out[threadIdx.x] = 0.0;
      return;
    }
  }
  batched_body<T, B>(P, D, J, r);   // D.pos_out is the second position buffer
}

// One serving job: its sizes and the device arrays of its own context.
struct JobView {
  int N, nchains, maxlen, capb, lpb, rk, todo, mc_first;
  int maxnb, maxw, maxinc, maxbpb;
  int64_t stepno;
  uint64_t seed;
  const Topo *topo;
  const int *chain_id, *bp3_ctx;
  const double *md_a, *md_cf, *md_cx;
  double *md_cr, *tK, *ele_coef, *lambdaD, *ele_cut, *box;
  double4 *pos, *pos_b, *vel, *acc, *move, *frc, *fbond;
  int *nb_ele, *n_ele, *nb_wca, *n_wca, *bp_tmp, *n_bp_tmp, *bp_off;
  int *bp_mp;
  double *bp_coef;
  int *inc, *n_inc;
  uint8_t *bp_status, *bp_status_mc;
  double *ene_bp, *for_bp;
  int *nt_cnt, *comp, *flags;
  int *seg_n, *seg, *has_exc;
  int *io;   // [0] steps done, [1] candidates, [2] rebuild pending, [3] rebuilds
};

// Job pool: workgroup j advances job j.  P carries the force field shared
// by the jobs (sizes are the job's own); D0 carries the shared tables.
template <typename T, int B>
__global__ void __launch_bounds__(B) k_jobs(Params P, Ptrs D0, const JobView *__restrict__ jobs, int xlds,
                                            int fplds) {
  const JobView &V = jobs[blockIdx.x];
  Params Pj = P;
  Pj.N = V.N;
  Pj.R = 1;
  Pj.nchains = V.nchains;
  Pj.maxlen = V.maxlen;
  Ptrs D = D0;
  D.topo = V.topo; D.chain_id = V.chain_id; D.bp3_ctx = V.bp3_ctx;
  D.md_a = V.md_a; D.md_cf = V.md_cf; D.md_cx = V.md_cx; D.md_cr = V.md_cr;
  D.tK = V.tK; D.ele_coef = V.ele_coef; D.lambdaD = V.lambdaD; D.ele_cut = V.ele_cut; D.box = V.box;
  D.pos = V.pos; D.vel = V.vel; D.acc = V.acc; D.move = V.move; D.frc = V.frc; D.fbond = V.fbond;
  D.pos_out = V.pos_b;
  D.maxnb = V.maxnb; D.maxw = V.maxw; D.maxinc = V.maxinc; D.maxbpb = V.maxbpb;
  D.nb_ele = V.nb_ele; D.n_ele = V.n_ele; D.nb_wca = V.nb_wca; D.n_wca = V.n_wca;
  D.bp_tmp = V.bp_tmp; D.n_bp_tmp = V.n_bp_tmp; D.bp_off = V.bp_off;
  D.bp_mp = V.bp_mp; D.bp_coef = V.bp_coef; D.inc = V.inc; D.n_inc = V.n_inc;
  D.bp_status = V.bp_status; D.bp_status_mc = V.bp_status_mc; D.ene_bp = V.ene_bp; D.for_bp = V.for_bp;
  D.nt_cnt = V.nt_cnt; D.comp = V.comp; D.flags = V.flags;
  SysIO J;
  J.todo = V.todo;
  J.mc_first = V.mc_first;
  J.capb = V.capb;
  J.lpb = V.lpb;
  J.rk = V.rk;
  J.stepno = V.stepno;
  J.seed = V.seed;
  J.seg_n = V.seg_n;
  J.seg = V.seg;
  J.has_exc = V.has_exc;
  J.pos_b = V.pos_b;
  J.xlds = xlds;
  J.fplds = fplds;
  J.nblds = 0;
  J.stlds = 0;
  J.mplds = 0;
  J.force_only = 0;
  J.energy_mode = 0;
  J.e_coef = 0.0;
  J.e_lam = 1.0;
  J.energy_out = nullptr;
  J.stepno_out = nullptr;
  J.mc_first_out = nullptr;
  J.done = V.io;
  J.nbp = V.io + 1;
  J.rebuild = V.io + 2;
  J.nreb = V.io + 3;
  batched_body<T, B>(Pj, D, J, 0);
}

}  // namespace sis
