// Phase kernels of the SIS MD step (one thread per bead, per candidate or
// per component).  Included by sis_gpu.hip.
#pragma once
#include "sis_device.h"
#include "sis_mc_wave.h"

namespace sis {

// ---------------------------------------------------------------------
// pbc_wrap (src/pbc.F90:196-226) for replicas with mask[r] != 0.
// pbc_wrap of the replicas being rebuilt (src/pbc.F90:196-226), and their
// electrostatic cutoff for the new lists (src/neighbor_list.F90:80-82)
__global__ void k_wrap(Params P, Ptrs D, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (mask && !mask[r]) return;
  if (i == 0) D.ele_cut[r] = rebuild_ele_cut(P, D, r);
  if (i >= P.N || !P.pbc) return;
  const double *box = D.box + 3 * r;
  double4 &p = D.pos[(size_t)r * P.N + i];
  double v[3] = {p.x, p.y, p.z};
  for (int k = 0; k < 3; k++) {
    double L = box[k];
    if (v[k] > L) v[k] = v[k] - L * ((double)(int)((v[k] - L) / L) + 1);
    else if (v[k] < 0.0) v[k] = v[k] + L * ((double)(int)((0.0 - v[k]) / L) + 1);
  }
  p.x = v[0]; p.y = v[1]; p.z = v[2];
}

// ---------------------------------------------------------------------
// Base-pair eligibility of replica-local beads i < j (0-based) and the
// triplet-table index: the pair rules of src/init_bp.F90:184-313 (models
// 3/4/5); with a secondary-structure file only its pairs are considered
// (lines 206-209).  Returns table index (0 = not eligible); *bpt receives
// the BPT type.
__device__ __forceinline__ int is_comp(int a, int b) {
  return (a == SQ_A && b == SQ_U) || (a == SQ_U && (b == SQ_A || b == SQ_G)) ||
         (a == SQ_G && (b == SQ_C || b == SQ_U)) || (a == SQ_C && b == SQ_G);
}
__device__ __forceinline__ int hkey(int w, int x, int u, int z, int y, int v) {
  if (x < y) return 1 + w + x * 4 + u * 16 + z * 64 + y * 256 + v * 1024;
  return 1 + v + y * 4 + z * 16 + u * 64 + x * 256 + w * 1024;
}
__device__ __forceinline__ int bp_eligible(const Params &P, const Ptrs &D, int i, int j, int *bpt) {
  *bpt = 0;
  if (D.ss_off) {
    const int a = i < j ? i : j, b = i < j ? j : i;
    int lo = D.ss_off[a], hi = D.ss_off[a + 1];
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (D.ss_partner[m] < b) lo = m + 1; else hi = m;
    }
    if (lo == D.ss_off[a + 1] || D.ss_partner[lo] != b) return 0;
  }
  Topo ti = D.topo[i], tj = D.topo[j];
  int li = ti.loc, lj = tj.loc, ni = ti.len, nj = tj.len;
  bool ic = topo_circ(ti), jc = topo_circ(tj);
  int ipre = li - 1, inxt = li + 1, jpre = lj - 1, jnxt = lj + 1;
  if (ic) { if (li == 1) ipre = ni; if (li == ni) inxt = 1; }
  else if (li == 1 || li == ni) return 0;
  if (jc) { if (lj == 1) jpre = nj; if (lj == nj) jnxt = 1; }
  else if (lj == 1 || lj == nj) return 0;
  auto sq = [&](const Topo &t, int l) { return topo_seq(D.topo[t.start + l - 1]); };
  int s_i = topo_seq(ti), s_j = topo_seq(tj);
  int s_ipre = sq(ti, ipre), s_inxt = sq(ti, inxt), s_jpre = sq(tj, jpre), s_jnxt = sq(tj, jnxt);
  int p1 = is_comp(s_ipre, s_jnxt), p2 = is_comp(s_inxt, s_jpre);
  if (!ic && ipre == 1) p1 = 0;
  if (!jc && jnxt == nj) p1 = 0;
  if (!ic && inxt == ni) p2 = 0;
  if (!jc && jpre == 1) p2 = 0;
  int L = P.bp_min_loop;
  if (ti.start == tj.start) {
    if (li + L >= lj) return 0;
    if (li + 1 + L >= lj - 1) p2 = 0;
    if (ic) {
      if (lj + L >= li + ni) return 0;
      if (lj + 1 + L >= li - 1 + ni) p1 = 0;
    }
  }
  if ((P.bp_model >= 3) && !p1 && !p2) return 0;
  int t = 0;
  if (s_i == SQ_G && s_j == SQ_C) t = 1; else if (s_i == SQ_C && s_j == SQ_G) t = 4;
  else if (s_i == SQ_A && s_j == SQ_U) t = 2; else if (s_i == SQ_U && s_j == SQ_A) t = 5;
  else if (s_i == SQ_G && s_j == SQ_U) t = 3; else if (s_i == SQ_U && s_j == SQ_G) t = 6;
  *bpt = t;
  if (P.bp_model != 5) return t;
  int key;
  if (!p1) key = hkey(SQ_U, s_i, s_inxt, SQ_U, s_j, s_jpre);
  else if (!p2) key = hkey(s_ipre, s_i, SQ_U, s_jnxt, s_j, SQ_U);
  else key = hkey(s_ipre, s_i, s_inxt, s_jnxt, s_j, s_jpre);
  return D.bp3_hash[key];
}

// Pair criteria of src/neighbor_list.F90:123-223 for replica-local beads
// a != b of replica r; appends to the owner's (a's) full lists.
struct SearchOut {
  int ne, nw, nb;
};
__device__ __forceinline__ void pair_test(const Params &P, const Ptrs &D, int r, int a, int b,
                                          const double4 &pa, const double4 &pb, double ele_cut2,
                                          SearchOut &o) {
  const int N = P.N;
  int lo = a < b ? a : b, hi = a < b ? b : a;
  const double4 &plo = a < b ? pa : pb;
  const double4 &phi = a < b ? pb : pa;
  double vx, vy, vz;
  pbc_vec(D.box + 3 * r, P.pbc, plo.x, plo.y, plo.z, phi.x, phi.y, phi.z, vx, vy, vz);
  double d2 = dot_nc(vx, vy, vz, vx, vy, vz);
  if (d2 > ele_cut2 && d2 > P.wca_nl_cut2 && d2 > P.bp_nl_cut2) return;
  Topo tl = D.topo[lo], th = D.topo[hi];
  bool same = tl.start == th.start;
  int il = tl.loc, jl = th.loc, n = tl.len;
  bool circ = topo_circ(tl);
  // the analysis lists take no circular wrap into their exclusions
  // (src/list_exv.F90:34-38, src/list_ele.F90:35-46)
  const bool nl_rule = P.list_rule == 0;
  // WCA
  if (topo_seq(tl) != SQ_D && topo_seq(th) != SQ_D && d2 <= P.wca_nl_cut2) {
    bool add = true;
    if (same) {
      if (il + 2 >= jl) add = false;
      else if (nl_rule && circ && jl + 2 >= il + n) add = false;
    }
    if (add) {
      if (o.nw < D.maxw) D.nb_wca[((size_t)r * D.maxw + o.nw) * N + a] = b;
      o.nw++;
    }
  }
  // Debye-Hueckel
  if (P.ele_on && d2 <= ele_cut2 && topo_charged(tl) && topo_charged(th)) {
    bool add = true;
    // list_ele tests the chain-local indices of every chain pair, the
    // neighbour list those of one chain only
    if (same || !nl_rule) {
      if (jl == il + 1 && P.ele_excl_bond) add = false;
      if (jl == il + 2 && P.ele_excl_angl) add = false;
    }
    if (same && nl_rule) {
      if (circ) {
        if (il == 1) {
          if (jl == n && P.ele_excl_bond) add = false;
          if (jl == n - 1 && P.ele_excl_angl) add = false;
        } else if (il == 2) {
          if (jl == n && P.ele_excl_angl) add = false;
        }
      }
    }
    if (add) {
      if (o.ne < D.maxnb) D.nb_ele[((size_t)r * D.maxnb + o.ne) * N + a] = b;
      o.ne++;
    }
  }
  // base-pair candidates, owned by the lower bead; list_bp starts jmp at
  // imp + bp_min_loop + 1 over the whole system (src/list_bp.F90:41-42).
  // Candidates beyond the neighbour cutoff are left out in both rules:
  // they can form no pair, so no energy, force or limit-MC choice sees them.
  if (a < b && P.bp_model > 0 && d2 <= P.bp_nl_cut2 && (nl_rule || b >= a + P.bp_min_loop + 1)) {
    int bt;
    if (bp_eligible(P, D, a, b, &bt) > 0) {
      if (o.nb < D.maxbpb) D.bp_tmp[((size_t)r * N + a) * D.maxbpb + o.nb] = b;
      o.nb++;
    }
  }
}

__device__ __forceinline__ void search_finish(const Params &P, const Ptrs &D, int r, int a,
                                              SearchOut &o) {
  size_t g = (size_t)r * P.N + a;
  D.n_ele[g] = o.ne;
  D.n_wca[g] = o.nw;
  int nb = o.nb < D.maxbpb ? o.nb : D.maxbpb;
  D.n_bp_tmp[g] = nb;
  // sort this bead's BP partners so the global list follows the
  // reference order (chain_i, chain_j, i, j)
  int *t = D.bp_tmp + g * D.maxbpb;
  for (int x = 1; x < nb; x++) {
    int v = t[x], y = x - 1;
    while (y >= 0 && t[y] > v) { t[y + 1] = t[y]; y--; }
    t[y + 1] = v;
  }
  if (o.ne > D.maxnb || o.nw > D.maxw || o.nb > D.maxbpb) {
    D.flags[0] = 1;
    atomicMax(&D.flags[2], o.ne);
    atomicMax(&D.flags[3], o.nw);
    atomicMax(&D.flags[4], o.nb);
  }
}

// All-pairs search through LDS tiles (small systems or small boxes).
template <int B>
__global__ void __launch_bounds__(B) k_search_brute(Params P, Ptrs D, const int *mask) {
  __shared__ double4 sp[B];
  int r = blockIdx.y;
  if (mask && !mask[r]) return;
  int a = blockIdx.x * B + threadIdx.x;
  const int N = P.N;
  bool act = a < N;
  double4 pa = act ? D.pos[(size_t)r * N + a] : double4{0, 0, 0, 0};
  double ec = D.ele_cut[r] + P.nl_margin;
  double ele_cut2 = P.ele_on ? ec * ec : 0.0;
  SearchOut o{0, 0, 0};
  for (int t0 = 0; t0 < N; t0 += B) {
    __syncthreads();
    if (t0 + threadIdx.x < N) sp[threadIdx.x] = D.pos[(size_t)r * N + t0 + threadIdx.x];
    __syncthreads();
    if (act) {
      int lim = min(B, N - t0);
      for (int k = 0; k < lim; k++) {
        int b = t0 + k;
        if (b == a) continue;
        pair_test(P, D, r, a, b, pa, sp[k], ele_cut2, o);
      }
    }
  }
  if (act) search_finish(P, D, r, a, o);
}

// Cell list: assignment / fill / per-cell sort / search.
__device__ __forceinline__ int cell_index(const Ptrs &D, int r, const double4 &p) {
  const double *box = D.box + 3 * r;
  int cx = (int)(p.x * D.ncell_x / box[0]), cy = (int)(p.y * D.ncell_y / box[1]),
      cz = (int)(p.z * D.ncell_z / box[2]);
  cx = min(max(cx, 0), D.ncell_x - 1);
  cy = min(max(cy, 0), D.ncell_y - 1);
  cz = min(max(cz, 0), D.ncell_z - 1);
  return (cz * D.ncell_y + cy) * D.ncell_x + cx;
}
__global__ void k_cell_count(Params P, Ptrs D, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r])) return;
  int c = cell_index(D, r, D.pos[(size_t)r * P.N + i]);
  D.cell_of[(size_t)r * P.N + i] = c;
  atomicAdd(&D.cell_cnt[(size_t)r * D.ncell + c], 1);
}
__global__ void k_cell_fill(Params P, Ptrs D, const int *mask, int *fill) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r])) return;
  size_t cg = (size_t)r * D.ncell + D.cell_of[(size_t)r * P.N + i];
  int slot = atomicAdd(&fill[cg], 1);
  D.cell_beads[D.cell_start[cg] + slot] = i;
}
__global__ void k_cell_sort(Params P, Ptrs D, const int *mask) {
  size_t cg = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cg >= (size_t)P.R * D.ncell) return;
  int r = (int)(cg / D.ncell);
  if (mask && !mask[r]) return;
  int s = D.cell_start[cg], e = D.cell_start[cg + 1];
  int *t = D.cell_beads;
  for (int x = s + 1; x < e; x++) {
    int v = t[x], y = x - 1;
    while (y >= s && t[y] > v) { t[y + 1] = t[y]; y--; }
    t[y + 1] = v;
  }
}
__global__ void k_search_cell(Params P, Ptrs D, const int *mask) {
  int a = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (a >= P.N || (mask && !mask[r])) return;
  const int N = P.N;
  double4 pa = D.pos[(size_t)r * N + a];
  double ec = D.ele_cut[r] + P.nl_margin;
  double ele_cut2 = P.ele_on ? ec * ec : 0.0;
  int c = D.cell_of[(size_t)r * N + a];
  int cx = c % D.ncell_x, cy = (c / D.ncell_x) % D.ncell_y, cz = c / (D.ncell_x * D.ncell_y);
  SearchOut o{0, 0, 0};
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int x = (cx + dx + D.ncell_x) % D.ncell_x, y = (cy + dy + D.ncell_y) % D.ncell_y,
            z = (cz + dz + D.ncell_z) % D.ncell_z;
        size_t cg = (size_t)r * D.ncell + (z * D.ncell_y + y) * D.ncell_x + x;
        int s = D.cell_start[cg], e = D.cell_start[cg + 1];
        for (int q = s; q < e; q++) {
          int b = D.cell_beads[q];
          if (b == a) continue;
          pair_test(P, D, r, a, b, pa, D.pos[(size_t)r * N + b], ele_cut2, o);
        }
      }
  search_finish(P, D, r, a, o);
}

// ---------------------------------------------------------------------
// Exclusive scan (three passes) over n ints.
template <int B>
__global__ void k_scan_blocks(const int *in, int *out, int *bsum, int64_t n) {
  __shared__ int s[B];
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  int v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < B; off <<= 1) {
    int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == B - 1) bsum[blockIdx.x] = s[B - 1];
}
__global__ void k_scan_bsum(int *bsum, int nb, int *total, int *out_end) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int acc = 0;
  for (int k = 0; k < nb; k++) { int v = bsum[k]; bsum[k] = acc; acc += v; }
  *total = acc;
  if (out_end) *out_end = acc;
}
template <int B>
__global__ void k_scan_add(int *out, const int *bsum, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
  if (i < n) out[i] += bsum[blockIdx.x];
}

// ---------------------------------------------------------------------
// Fill the BP candidate list (src/neighbor_list.F90:183-223) and scatter
// incidences.  bp_mp holds replica-local 1-based ids as the reference.
__global__ void k_bp_fill(Params P, Ptrs D, const int *mask) {
  int a = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (a >= P.N || (mask && !mask[r])) return;
  const int N = P.N;
  size_t g = (size_t)r * N + a;
  int nb = D.n_bp_tmp[g], off = D.bp_off[g];
  Topo ti = D.topo[a];
  for (int k = 0; k < nb; k++) {
    int b = D.bp_tmp[g * D.maxbpb + k];
    int gb = off + k;
    if (gb >= D.maxbp) { D.flags[0] = 1; continue; }
    Topo tj = D.topo[b];
    int i = ti.loc, j = tj.loc, imp = a + 1, jmp = b + 1;
    int mp[7];
    mp[0] = imp;
    mp[1] = jmp;
    if (P.list_rule == 0) {
      mp[2] = (i == 1) ? imp + ti.len - 1 : imp - 1;
      mp[3] = (j == 1) ? jmp + tj.len - 1 : jmp - 1;
      mp[4] = (i == ti.len) ? imp + 1 - ti.len : imp + 1;
      // the reference compares j with the length of chain i here
      // (src/neighbor_list.F90:212); kept unless it leaves the system
      int m6 = (j == ti.len) ? jmp + 1 - tj.len : jmp + 1;
      if (P.fix_h4 || m6 < 1 || m6 > N) m6 = (j == tj.len) ? jmp + 1 - tj.len : jmp + 1;
      mp[5] = m6;
    } else {
      // list_bp wraps the neighbours of circular chains only and takes
      // each bead's own chain length (src/list_bp.F90:60-82)
      const bool ic = topo_circ(ti), jc = topo_circ(tj);
      mp[2] = (ic && i == 1) ? imp + ti.len - 1 : imp - 1;
      mp[3] = (jc && j == 1) ? jmp + tj.len - 1 : jmp - 1;
      mp[4] = (ic && i == ti.len) ? imp + 1 - ti.len : imp + 1;
      mp[5] = (jc && j == tj.len) ? jmp + 1 - tj.len : jmp + 1;
    }
    int bt;
    int h = bp_eligible(P, D, a, b, &bt);
    mp[6] = bt;
    int *dst = D.bp_mp + (size_t)gb * 8;
    for (int q = 0; q < 7; q++) dst[q] = mp[q];
    dst[7] = r;
    // the triplet table exists for model 5 only, other models read zeros
    // (src/init_bp.F90:39, src/neighbor_list.F90:220-221)
    D.bp_coef[2 * (size_t)gb] = P.bp_model == 5 ? D.bp3_dH[h] : 0.0;
    D.bp_coef[2 * (size_t)gb + 1] = P.bp_model == 5 ? D.bp3_dS[h] : 0.0;
    // reference list order: (chain_i, chain_j, i, j) for the neighbour list
    // (src/neighbor_list.F90:99-114), (imp, jmp) for list_bp
    // (src/list_bp.F90:41-42)
    if (P.list_rule == 0)
      D.bp_key[gb] = (((unsigned long long)(r * (int64_t)P.nchains + D.chain_id[a]) * P.nchains +
                       D.chain_id[b]) * P.maxlen + (i - 1)) * P.maxlen + (j - 1);
    else
      D.bp_key[gb] = ((unsigned long long)r * N + a) * N + b;
    D.bp_val[gb] = gb;
  }
}
// gather the candidate records into key order
__global__ void k_bp_permute(Ptrs D, int total, const int *order, int *mp_out, double *coef_out) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  int o = order[p];
  for (int q = 0; q < 8; q++) mp_out[(size_t)p * 8 + q] = D.bp_mp[(size_t)o * 8 + q];
  coef_out[2 * (size_t)p] = D.bp_coef[2 * (size_t)o];
  coef_out[2 * (size_t)p + 1] = D.bp_coef[2 * (size_t)o + 1];
}
// incidences of each candidate on its six beads
__global__ void k_bp_inc(Params P, Ptrs D, int total) {
  int gb = blockIdx.x * blockDim.x + threadIdx.x;
  if (gb >= total) return;
  const int *mp = D.bp_mp + (size_t)gb * 8;
  int r = mp[7];
  const int N = P.N;
  for (int q = 0; q < 6; q++) {
    size_t gq = (size_t)r * N + (mp[q] - 1);
    int slot = atomicAdd(&D.n_inc[gq], 1);
    if (slot < D.maxinc) D.inc[((size_t)r * D.maxinc + slot) * N + (mp[q] - 1)] = gb * 8 + q;
    else { D.flags[0] = 1; atomicMax(&D.flags[6], slot + 1); }
  }
}
// deterministic incidence order: sort each bead's list by (candidate, role)
__global__ void k_inc_sort(Params P, Ptrs D, const int *mask) {
  int a = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (a >= P.N || (mask && !mask[r])) return;
  size_t g = (size_t)r * P.N + a;
  int n = min(D.n_inc[g], D.maxinc);
  size_t base = (size_t)r * D.maxinc * P.N + a, st = P.N;
  for (int x = 1; x < n; x++) {
    int v = D.inc[base + x * st], y = x - 1;
    while (y >= 0 && D.inc[base + y * st] > v) { D.inc[base + (y + 1) * st] = D.inc[base + y * st]; y--; }
    D.inc[base + (y + 1) * st] = v;
  }
  D.nt_cnt[g] = -P.max_bp_per_nt;
}

// per-nucleotide formed-pair counter, reset before each BP evaluation
__global__ void k_reset_cnt(Params P, Ptrs D, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r])) return;
  D.nt_cnt[(size_t)r * P.N + i] = -P.max_bp_per_nt;
}

__device__ __forceinline__ int uf_find(int *p, int x) {
  while (true) {
    int q = p[x];
    if (q == x) return x;
    int qq = p[q];
    if (qq != q) atomicCAS(&p[x], q, qq);
    x = q;
  }
}

// ---------------------------------------------------------------------
// nbp_total < 0: the count of the current lists, read from the device (the
// end of the candidates' scan; graph replays launch over the capacity)
__device__ __forceinline__ int bp_count(const Params &P, const Ptrs &D, int nbp_total) {
  return nbp_total >= 0 ? nbp_total : D.bp_off[(size_t)P.R * P.N];
}

// BP evaluation per candidate (src/force_bp_limit_triplet.F90:85-256 and
// src/force_bp_triplet.F90).  mc[r]: MC step for replica r.  ecT (may be
// null): temperature override per replica (energy at another condition).
template <typename T>
__global__ void k_bp_eval(Params P, Ptrs D, int nbp_total, const int *mc, const double *ecT,
                          const int *mask, int do_union) {
  int gb = blockIdx.x * blockDim.x + threadIdx.x;
  if (gb >= bp_count(P, D, nbp_total)) return;
  const int *mp = D.bp_mp + (size_t)gb * 8;
  int r = mp[7];
  if (mask && !mask[r]) return;
  const int N = P.N;
  D.bp_status[gb] = 0;
  D.ene_bp[gb] = 0.0;
  if (!mc[r] && P.nstep_bp_mc > 0 && !D.bp_status_mc[gb]) return;
  double tK = ecT ? ecT[r] : D.tK[r];
  // model 5: dG = dH - T dS of the triplet; models 1/3: U0 exp(-u) alone
  // (src/force_bp.F90:203)
  double dG = 1.0;
  if (P.bp_model == 5) {
    dG = D.bp_coef[2 * (size_t)gb] - tK * D.bp_coef[2 * (size_t)gb + 1];
    if (dG >= 0.0) return;
  }
  const double *box = D.box + 3 * r;
  const double4 *x = D.pos + (size_t)r * N;
  double4 p1 = x[mp[0] - 1], p2 = x[mp[1] - 1];
  V3<double> v12 = pvec(P, box, p1, p2), v13 = pvec(P, box, p1, x[mp[2] - 1]),
             v15 = pvec(P, box, p1, x[mp[4] - 1]), v42 = pvec(P, box, x[mp[3] - 1], p2),
             v62 = pvec(P, box, x[mp[5] - 1], p2);
  const BpPara &bp = D.bpp[mp[6]];
  T u;
  V3<T> fb[6];
  if (!bp_geom<T>(bp, v12, v13, v15, v42, v62, u, fb)) return;
  double e = bp.U0 * dG * exp(-(double)u);
  if (e <= P.bp_cutoff_energy) {
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
    // performance-form MC: join the pair's nucleotides into one conflict
    // component (union-find, larger root linked under the smaller)
    if (do_union && mc[r]) {
      int *pc = D.comp + (size_t)r * N;
      int ua = mp[0] - 1, ub = mp[1] - 1;
      while (true) {
        ua = uf_find(pc, ua);
        ub = uf_find(pc, ub);
        if (ua == ub) break;
        if (ua > ub) { int t = ua; ua = ub; ub = t; }
        if (atomicCAS(&pc[ub], ub, ua) == ub) break;
      }
    }
  }
}

// ---------------------------------------------------------------------
// Limit MC, parity form: the serial algorithm of
// src/force_bp_limit_triplet.F90:259-330 on the injected MT words of the
// replica, one thread.  Consumes words exactly as genrand_double3 /
// genrand_double1 would.
__global__ void k_mc_serial(Params P, Ptrs D, const int *mc, const double *ecT, int *ntlist,
                            int *bpseq, const int *mask) {
  if (threadIdx.x != 0) return;
  int r = blockIdx.x;
  if (mask && !mask[r]) return;
  if (!mc[r]) return;
  const int N = P.N;
  double tK = ecT ? ecT[r] : D.tK[r];
  double beta = mc_beta(P, tK, ecT == nullptr);
  int b0 = D.bp_off[(size_t)r * N], nb = D.bp_off[(size_t)(r + 1) * N] - b0;
  const uint32_t *w = D.tape + D.tape_off[r];
  int64_t pos = D.tape_pos[r], lim = D.tape_off[r + 1] - D.tape_off[r];
  int *cnt = D.nt_cnt + (size_t)r * N;
  int *nl = ntlist + (size_t)r * (N + 1);
  int *bs = bpseq + b0 + r;
  for (;;) {
    int nnt = 0;
    for (int k = 0; k < N; k++) if (cnt[k] > 0) nl[++nnt] = k + 1;
    if (nnt == 0) break;
    if (pos + 2 > lim) { D.flags[1] = 1; break; }
    double rnd = mt_d3(w[pos], w[pos + 1]); pos += 2;
    int ntd = nl[(int)ceil(rnd * nnt)];
    int ns = 0;
    for (int k = 0; k < nb; k++) {
      int gb = b0 + k;
      if (D.bp_status[gb]) {
        const int *mp = D.bp_mp + (size_t)gb * 8;
        if (mp[0] == ntd || mp[1] == ntd) bs[++ns] = gb;
      }
    }
    for (int k = 1; k <= ns; k++) {
      if (pos + 2 > lim) { D.flags[1] = 1; break; }
      rnd = mt_d3(w[pos], w[pos + 1]); pos += 2;
      int sw = (int)ceil(rnd * ns);
      int sv = bs[k]; bs[k] = bs[sw]; bs[sw] = sv;
    }
    if (ns == 0) { D.flags[1] = 2; break; }
    int del = bs[1];
    for (int k = 2; k <= ns; k++) {
      int jb = bs[k];
      double ratio = exp((D.ene_bp[jb] - D.ene_bp[del]) * beta);
      if (pos + 2 > lim) { D.flags[1] = 1; break; }
      rnd = mt_d1(w[pos], w[pos + 1]); pos += 2;
      if (rnd < ratio) del = jb;
    }
    D.bp_status[del] = 0;
    const int *mp = D.bp_mp + (size_t)del * 8;
    cnt[mp[0] - 1] -= 1;
    cnt[mp[1] - 1] -= 1;
  }
  D.tape_pos[r] = pos;
}

// Limit MC, performance form.  The excess-removal process of the reference
// decomposes over connected components of the formed-pair graph: a step
// only changes counts inside the component of the chosen nucleotide, and
// the uniform choice among all excess nucleotides restricted to one
// component is uniform within it.  Each component is therefore resolved
// independently by one thread with its own counter-RNG stream.
__global__ void k_uf_init(Params P, Ptrs D, int *cnt_exc, int *fill, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r])) return;
  size_t g = (size_t)r * P.N + i;
  D.comp[g] = i;
  cnt_exc[g] = 0;
  fill[g] = 0;
}
__global__ void k_uf_union(Params P, Ptrs D, int nbp_total, const int *mc, const int *mask) {
  int gb = blockIdx.x * blockDim.x + threadIdx.x;
  if (gb >= nbp_total || !D.bp_status[gb]) return;
  const int *mp = D.bp_mp + (size_t)gb * 8;
  int r = mp[7];
  if ((mask && !mask[r]) || !mc[r]) return;
  int *p = D.comp + (size_t)r * P.N;
  int a = mp[0] - 1, b = mp[1] - 1;
  while (true) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return;
    if (a > b) { int t = a; a = b; b = t; }
    if (atomicCAS(&p[b], b, a) == b) return;  // link larger root under smaller
  }
}
__global__ void k_uf_count(Params P, Ptrs D, int *cnt_exc, const int *mc, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r]) || !mc[r]) return;
  size_t g = (size_t)r * P.N + i;
  if (D.nt_cnt[g] > 0) {
    int root = uf_find(D.comp + (size_t)r * P.N, i);
    atomicAdd(&cnt_exc[(size_t)r * P.N + root], 1);
  }
}
__global__ void k_uf_fill(Params P, Ptrs D, const int *exc_off, int *fill, int *exc_list,
                          const int *mc, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r]) || !mc[r]) return;
  size_t g = (size_t)r * P.N + i;
  if (D.nt_cnt[g] > 0) {
    int root = uf_find(D.comp + (size_t)r * P.N, i);
    size_t gr = (size_t)r * P.N + root;
    int slot = atomicAdd(&fill[gr], 1);
    exc_list[exc_off[gr] + slot] = i;
  }
}
__global__ void k_mc_comp(Params P, Ptrs D, const int *exc_off, const int *cnt_exc,
                          int *exc_list, const int *mc, const double *ecT, uint64_t seed,
                          int64_t step, const int *mask) {
  int root = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (root >= P.N || (mask && !mask[r]) || !mc[r]) return;
  const int N = P.N;
  size_t gr = (size_t)r * N + root;
  int ne = cnt_exc[gr];
  if (ne == 0) return;
  int *E = exc_list + exc_off[gr];
  for (int x = 1; x < ne; x++) {
    int v = E[x], y = x - 1;
    while (y >= 0 && E[y] > v) { E[y + 1] = E[y]; y--; }
    E[y + 1] = v;
  }
  double tK = ecT ? ecT[r] : D.tK[r];
  double beta = mc_beta(P, tK, ecT == nullptr);
  int *cnt = D.nt_cnt + (size_t)r * N;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32) ^ (0x5bd1e995u * (uint32_t)(r + 1));
  uint32_t draw = 0;
  auto uni = [&]() {
    U4 o = philox(U4{(uint32_t)step, (uint32_t)(step >> 32) ^ 0x4d43u, (uint32_t)root, draw++}, k0, k1);
    return u01(o.x, o.y);
  };
  const int CAP = 64;
  int seqv[CAP];
  size_t st = N;
  for (;;) {
    int nact = 0;
    for (int k = 0; k < ne; k++) nact += cnt[E[k]] > 0;
    if (nact == 0) break;
    int pick = (int)ceil(uni() * nact), ntd = -1;
    for (int k = 0, c = 0; k < ne; k++)
      if (cnt[E[k]] > 0 && ++c == pick) { ntd = E[k]; break; }
    if (ntd < 0) break;
    int ni = min(D.n_inc[(size_t)r * N + ntd], D.maxinc), ns = 0;
    const int *inc = D.inc + (size_t)r * D.maxinc * N + ntd;
    for (int k = 0; k < ni && ns < CAP; k++) {
      int e = inc[k * st], gb = e >> 3, role = e & 7;
      if (role <= 1 && D.bp_status[gb]) seqv[ns++] = gb;
    }
    for (int k = 0; k < ns; k++) {
      int sw = (int)ceil(uni() * ns) - 1;
      int sv = seqv[k]; seqv[k] = seqv[sw]; seqv[sw] = sv;
    }
    if (ns == 0) { D.flags[1] = 2; break; }
    int del = seqv[0];
    for (int k = 1; k < ns; k++) {
      int jb = seqv[k];
      double ratio = exp((D.ene_bp[jb] - D.ene_bp[del]) * beta);
      if (uni() < ratio) del = jb;
    }
    D.bp_status[del] = 0;
    const int *mp = D.bp_mp + (size_t)del * 8;
    cnt[mp[0] - 1] -= 1;
    cnt[mp[1] - 1] -= 1;
  }
}

// Performance-form components: union-find roots, pairs bucketed per root
// (capacity SEGCAP per component), one wavefront per component with excess.
constexpr int SEGCAP = 256;
__global__ void k_comp_init(Params P, Ptrs D, int *seg_n, int *has_exc, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r])) return;
  size_t g = (size_t)r * P.N + i;
  D.comp[g] = i;
  seg_n[g] = 0;
  has_exc[g] = 0;
}
// the per-step state of the performance-form MC in one pass: formed-pair
// counts (k_reset_cnt), union-find parents and buckets (k_comp_init), and
// the list of components with excess (its count in exc_list[0])
__global__ void k_mc_reset(Params P, Ptrs D, int *seg_n, int *has_exc, int *exc_list, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i == 0 && r == 0) exc_list[0] = 0;
  if (i >= P.N || (mask && !mask[r])) return;
  size_t g = (size_t)r * P.N + i;
  D.nt_cnt[g] = -P.max_bp_per_nt;
  D.comp[g] = i;
  seg_n[g] = 0;
  has_exc[g] = 0;
}
__global__ void k_seg_fill(Params P, Ptrs D, int nbp_total, int *seg_n, int *seg, int *has_exc,
                           const int *mc, const int *mask, int *exc_list) {
  int gb = blockIdx.x * blockDim.x + threadIdx.x;
  if (gb >= bp_count(P, D, nbp_total) || !D.bp_status[gb]) return;
  const int *mp = D.bp_mp + (size_t)gb * 8;
  int r = mp[7];
  if ((mask && !mask[r]) || !mc[r]) return;
  const int N = P.N;
  int a = mp[0] - 1, b = mp[1] - 1;
  // the counts (final since k_bp_eval) are loaded before the dependent
  // chain of the root walk and the bucket atomic, not after it
  const bool exc = D.nt_cnt[(size_t)r * N + a] > 0 || D.nt_cnt[(size_t)r * N + b] > 0;
  int root = uf_find(D.comp + (size_t)r * N, a);
  size_t gr = (size_t)r * N + root;
  int slot = atomicAdd(&seg_n[gr], 1);
  if (slot < SEGCAP) seg[gr * SEGCAP + slot] = gb;
  if (exc) {
    // listed once (exc_list: [0] count, then replica * N + root)
    if (atomicExch(&has_exc[gr], 1) == 0 && exc_list) exc_list[1 + atomicAdd(&exc_list[0], 1)] = (int)gr;
  }
}
// serial fallback for a component larger than one wave holds: the same
// algorithm on global memory by one lane
// The arrays it touches come by value so that the callers' parameter
// structs stay in registers.
struct McArrays {
  const int *bp_mp;
  uint8_t *bp_status;
  const double *ene_bp;
  int *nt_cnt, *flags;
};
__device__ __noinline__ void mc_component_serial(const McArrays D, const int N, int r, int rk, int root, int *seg,
                                                 int n, double beta, uint64_t seed, int64_t step) {
  int *cnt = D.nt_cnt + (size_t)r * N;
  for (int t = 0; t < n; t++) {
    const int *mp = D.bp_mp + (size_t)seg[t] * 8;
    if (mp[0] < 1 || mp[0] > N || mp[1] < 1 || mp[1] > N) { D.flags[7] = 1; return; }
  }
  for (int x = 1; x < n; x++) {
    int v = seg[x], y = x - 1;
    while (y >= 0 && seg[y] > v) { seg[y + 1] = seg[y]; y--; }
    seg[y + 1] = v;
  }
  uint32_t draw = 0x10000u;
  int sq[64];
  for (int iter = 0; iter <= n; iter++) {
    // active nucleotides = endpoints with count > 0, ascending id
    int nact = 0, prev = -1;
    while (true) {
      int nxt = 0x7fffffff;
      for (int t = 0; t < n; t++) {
        const int *mp = D.bp_mp + (size_t)seg[t] * 8;
        int a = mp[0] - 1, b = mp[1] - 1;
        if (a > prev && a < nxt && cnt[a] > 0) nxt = a;
        if (b > prev && b < nxt && cnt[b] > 0) nxt = b;
      }
      if (nxt == 0x7fffffff) break;
      nact++;
      prev = nxt;
    }
    if (nact == 0) return;
    int pick = (int)ceil(philox_u(seed, rk, step, root, draw++) * nact), ntd = -1;
    prev = -1;
    for (int c = 0; c < pick; c++) {
      int nxt = 0x7fffffff;
      for (int t = 0; t < n; t++) {
        const int *mp = D.bp_mp + (size_t)seg[t] * 8;
        int a = mp[0] - 1, b = mp[1] - 1;
        if (a > prev && a < nxt && cnt[a] > 0) nxt = a;
        if (b > prev && b < nxt && cnt[b] > 0) nxt = b;
      }
      prev = nxt;
    }
    ntd = prev;
    int ns = 0;
    for (int t = 0; t < n && ns < 64; t++) {
      int gb = seg[t];
      const int *mp = D.bp_mp + (size_t)gb * 8;
      if (D.bp_status[gb] && (mp[0] - 1 == ntd || mp[1] - 1 == ntd)) sq[ns++] = gb;
    }
    if (ns == 0) { D.flags[1] = 2; return; }
    for (int i = 0; i < ns; i++) {
      int sw = (int)ceil(philox_u(seed, rk, step, root, draw++) * ns) - 1;
      int t = sq[i]; sq[i] = sq[sw]; sq[sw] = t;
    }
    int del = sq[0];
    for (int i = 1; i < ns; i++) {
      double ratio = exp((D.ene_bp[sq[i]] - D.ene_bp[del]) * beta);
      if (philox_u(seed, rk, step, root, draw++) < ratio) del = sq[i];
    }
    D.bp_status[del] = 0;
    const int *mp = D.bp_mp + (size_t)del * 8;
    cnt[mp[0] - 1] -= 1;
    cnt[mp[1] - 1] -= 1;
  }
}
__device__ __forceinline__ McArrays mc_arrays(const Ptrs &D) {
  return McArrays{D.bp_mp, D.bp_status, D.ene_bp, D.nt_cnt, D.flags};
}

// Scratch (ints) that mc_component_huge needs for a replica with nc
// candidates and N beads.
__device__ __forceinline__ size_t mc_huge_ints(int nc, int N) {
  const size_t pc = (size_t)((nc + 1) & ~1);
  const size_t nw = (size_t)(N + 31) / 32;
  return 8 * pc + ((nw + (size_t)N + 1) & ~(size_t)1) + 2 * (size_t)((N + 63) / 64);
}

// A component of any size (those above the bucket capacity SEGCAP): one
// wavefront, the state of the sweep in global scratch.  The sweep is that
// of mc_component_wave (an excess nucleotide picked uniformly in ascending
// id order, its formed pairs in list order, the reference's shuffle and
// Metropolis sweep, one deletion per iteration), on its own draws:
//   - the component's formed pairs come from a scan of the replica's
//     candidates [g0, g1) in order (list order kept), keeping those whose
//     union-find root is `root`;
//   - its nucleotides are listed in ascending id order through a bitmap
//     over bead ids;
//   - active nucleotides are 64-bit masks, updated on each deletion;
//   - each iteration's sequence is a ballot scan over the pairs, staged
//     through sseq (64 ints of LDS).
// comp is the replica's union-find array; scr holds cap ints.  Returns
// false if cap is too small.  Must be called by all 64 lanes of a wave
// with uniform arguments.
__device__ __noinline__ bool mc_component_huge(const McArrays D, const int N, int r, int rk, int root, int *comp,
                                               int g0, int g1, double beta, uint64_t seed, int64_t step, int *scr,
                                               size_t cap, int *sseq) {
  const int L = lane_id();
  root = __builtin_amdgcn_readfirstlane(root);
  const int nc = g1 - g0;
  if (mc_huge_ints(nc, N) > cap) return false;
  const size_t pc = (size_t)((nc + 1) & ~1);
  const int nw = (N + 31) / 32;
  int *hp = scr, *ha = hp + pc, *hb = ha + pc, *hua = hb + pc, *hub = hua + pc, *hal = hub + pc;
  double *he = reinterpret_cast<double *>(hal + pc);
  uint32_t *bm = reinterpret_cast<uint32_t *>(he + pc);
  int *hn = reinterpret_cast<int *>(bm + nw);
  unsigned long long *am = reinterpret_cast<unsigned long long *>(hn + ((nw + N + 1) & ~1) - nw);
  int *cnt = D.nt_cnt + (size_t)r * N;
  // ---- the component's formed pairs, in candidate order
  int n = 0;
  bool bad = false;
  for (int g = g0; g < g1; g += 64) {
    const int gb = g + L;
    bool in = false;
    int a = 0, b = 0;
    if (gb < g1 && D.bp_status[gb]) {
      const int *mp = D.bp_mp + (size_t)gb * 8;
      a = mp[0] - 1;
      b = mp[1] - 1;
      if (a < 0 || a >= N || b < 0 || b >= N) bad = true;
      else in = uf_find(comp, a) == root;
    }
    const unsigned long long m = __ballot(in);
    if (in) {
      const int t = n + below(m);
      hp[t] = gb; ha[t] = a; hb[t] = b; hua[t] = 0; hub[t] = 0; hal[t] = 1;
      he[t] = D.ene_bp[gb];
    }
    n += __popcll(m);
  }
  if (__ballot(bad)) {
    if (L == 0) D.flags[7] = 1;
    return true;
  }
  // ---- its nucleotides in ascending id order
  for (int w = L; w < nw; w += 64) bm[w] = 0u;
  __threadfence();
  for (int t = L; t < n; t += 64) {
    atomicOr(&bm[ha[t] >> 5], 1u << (ha[t] & 31));
    atomicOr(&bm[hb[t] >> 5], 1u << (hb[t] & 31));
  }
  __threadfence();
  int U = 0;
  for (int w0 = 0; w0 < nw; w0 += 64) {
    const int w = w0 + L;
    const uint32_t bits = w < nw ? bm[w] : 0u;
    const int c = __popc(bits), inc = wave_incl_scan(c);
    int k = U + inc - c;
    for (uint32_t x = bits; x; x &= x - 1) hn[k++] = 32 * w + __ffs(x) - 1;
    U += bcast(inc, 63);
  }
  __threadfence();
  // node index of each endpoint, then the active masks
  for (int t = L; t < n; t += 64) {
    int v = ha[t], lo = 0, hi = U;
    while (lo < hi) { const int mid = (lo + hi) >> 1; if (hn[mid] < v) lo = mid + 1; else hi = mid; }
    hua[t] = lo;
    v = hb[t]; lo = 0; hi = U;
    while (lo < hi) { const int mid = (lo + hi) >> 1; if (hn[mid] < v) lo = mid + 1; else hi = mid; }
    hub[t] = lo;
  }
  const int nuw = (U + 63) / 64;
  for (int u0 = 0; u0 < U; u0 += 64) {
    const int u = u0 + L;
    const unsigned long long m = __ballot(u < U && cnt[hn[u < U ? u : 0]] > 0);
    if (L == 0) am[u0 >> 6] = m;
  }
  __threadfence();
  // ---- the sweep: one deletion per iteration
  uint32_t draw = 0;
  for (int iter = 0; iter <= n; iter++) {
    int nact = 0;
    for (int w0 = 0; w0 < nuw; w0 += 64) {
      const int w = w0 + L;
      nact += bcast(wave_incl_scan(w < nuw ? __popcll(am[w]) : 0), 63);
    }
    if (nact == 0) break;
    const int pick = __builtin_amdgcn_readfirstlane((int)ceil(philox_u(seed, rk, step, root, draw) * nact));
    int ntu = -1, acc = 0;
    for (int w0 = 0; w0 < nuw && ntu < 0; w0 += 64) {
      const int w = w0 + L;
      const unsigned long long m = w < nuw ? am[w] : 0ull;
      const int c = __popcll(m), inc = wave_incl_scan(c), tot = bcast(inc, 63);
      if (acc + tot >= pick) {
        const int want = pick - acc - 1;
        const unsigned long long sel = __ballot(c > 0 && want >= inc - c && want < inc);
        const int ln = __ffsll((long long)sel) - 1;
        if (ln < 0) break;
        const unsigned long long mw = bcast64(m, ln);
        const int k = want - (bcast(inc, ln) - bcast(c, ln));
        const unsigned long long hit = __ballot(((mw >> L) & 1ull) && below(mw) == k);
        ntu = 64 * (w0 + ln) + __ffsll((long long)hit) - 1;
      }
      acc += tot;
    }
    ntu = __builtin_amdgcn_readfirstlane(ntu);
    if (ntu < 0 || ntu >= U) {   // masks and counts disagree: corrupted state
      if (L == 0) D.flags[7] = 1;
      return true;
    }
    const int ntd = __builtin_amdgcn_readfirstlane(hn[ntu]);
    // formed pairs of ntd in list order (the first 64)
    int ns = 0;
    for (int t0 = 0; t0 < n && ns < 64; t0 += 64) {
      const int t = t0 + L;
      const bool hit = t < n && hal[t] && (ha[t] == ntd || hb[t] == ntd);
      const unsigned long long m = __ballot(hit);
      if (hit) {
        const int k = ns + below(m);
        if (k < 64) sseq[k] = t;
      }
      ns += __popcll(m);
    }
    if (ns == 0) {
      if (L == 0) D.flags[1] = 2;
      return true;
    }
    if (ns > 64) ns = 64;
    wave_lds_fence();
    int seqv = L < ns ? sseq[L] : 0;
    // draws after the pick: shuffle 1..ns, Metropolis ns+1..2ns-1
    const int swl = L < ns ? (int)ceil(philox_u(seed, rk, step, root, draw + 1u + (uint32_t)L) * ns) - 1 : 0;
    const double lgl = (L >= 1 && L < ns) ? log(philox_u(seed, rk, step, root, draw + (uint32_t)(ns + L))) : 0.0;
    for (int i = 0; i < ns; i++) {
      const int sw = bcast(swl, i);
      const int ti = bcast(seqv, i), ts = bcast(seqv, sw);
      seqv = setlane(seqv, ts, i);
      seqv = setlane(seqv, ti, sw);
    }
    const double ev = L < ns ? he[seqv] : 0.0;
    int del = bcast(seqv, 0);
    double edel = bcast(ev, 0);
    for (int i = 1; i < ns; i++) {
      const double ej = bcast(ev, i);
      if (bcast(lgl, i) < (ej - edel) * beta) { del = bcast(seqv, i); edel = ej; }
    }
    draw += 2u * (uint32_t)ns;
    if (L == 0) {
      hal[del] = 0;
      D.bp_status[hp[del]] = 0;
      if (atomicSub(&cnt[ha[del]], 1) <= 1) atomicAnd(&am[hua[del] >> 6], ~(1ull << (hua[del] & 63)));
      if (atomicSub(&cnt[hb[del]], 1) <= 1) atomicAnd(&am[hub[del] >> 6], ~(1ull << (hub[del] & 63)));
    }
    __threadfence();
  }
  return true;
}

// Per-phase form: components above SEGCAP pairs, one wavefront per
// replica after k_mc_wave (which marks the replicas that have them).
__global__ void __launch_bounds__(64) k_mc_huge(Params P, Ptrs D, const int *seg_n, int *seg, const int *has_exc,
                                                int *huge, const double *ecT, uint64_t seed, int64_t step0,
                                                const int64_t *step_dev) {
  const int r = blockIdx.x, L = threadIdx.x, N = P.N;
  if (!huge[r]) return;
  const int64_t step = step_dev ? *step_dev : step0;   // graph replays read the step from the device
  __shared__ int sseq[64];
  const double tK = ecT ? ecT[r] : D.tK[r];
  const double beta = mc_beta(P, tK, ecT == nullptr);
  const int g0 = D.bp_off[(size_t)r * N], g1 = D.bp_off[(size_t)(r + 1) * N];
  int *scr = seg + (size_t)r * N * SEGCAP;
  for (int u0 = 0; u0 < N; u0 += 64) {
    const int root = u0 + L;
    bool h = false;
    if (root < N) {
      const size_t gr = (size_t)r * N + root;
      h = has_exc[gr] && seg_n[gr] > SEGCAP;
    }
    for (unsigned long long m = __ballot(h); m; m &= m - 1) {
      const int rt = bcast(root, __ffsll((long long)m) - 1);
      if (!mc_component_huge(mc_arrays(D), N, r, r, rt, D.comp + (size_t)r * N, g0, g1, beta, seed, step, scr,
                             (size_t)N * SEGCAP, sseq))
        if (L == 0) D.flags[5] = 1;
    }
  }
  if (L == 0) huge[r] = 0;
}
// one conflict component (root of replica r) resolved by a 64-thread
// block: one lane, one wavefront, or the serial fallback; components above
// the bucket capacity are flagged for k_mc_huge
__device__ __forceinline__ void mc_one_component(const Params &P, const Ptrs &D, int r, int root, int *seg_n,
                                                 int *seg, const double *ecT, uint64_t seed, int64_t step,
                                                 int *huge, int *ws) {
  size_t gr = (size_t)r * P.N + root;
  int n = seg_n[gr];
  double tK = ecT ? ecT[r] : D.tK[r];
  double beta = mc_beta(P, tK, ecT == nullptr);
  if (n > SEGCAP) {   // above the bucket capacity: k_mc_huge takes it
    if (threadIdx.x == 0) huge[r] = 1;
    return;
  }
  int *sg = seg + gr * SEGCAP;
  if (n <= MCL_PAIRS) {   // small component: one lane, same result
    if (threadIdx.x == 0 && !mc_component_lane(D, P.N, r, r, root, sg, n, beta, seed, step))
      mc_component_serial(mc_arrays(D), P.N, r, r, root, sg, n, beta, seed, step);
    return;
  }
  const bool ok = n <= 64 * MCW_ITEMS ? mc_component_wave_any(P, D, r, r, root, sg, n, beta, seed, step, ws)
                                      : mc_component_wave<MCB_ITEMS, MCB_NODES>(P, D, r, r, root, sg, n, beta, seed, step, ws);
  if (!ok && threadIdx.x == 0) mc_component_serial(mc_arrays(D), P.N, r, r, root, sg, n, beta, seed, step);
}

__global__ void __launch_bounds__(64) k_mc_wave(Params P, Ptrs D, int *seg_n, int *seg,
                                                const int *has_exc, const int *mc,
                                                const double *ecT, uint64_t seed, int64_t step0,
                                                const int *mask, int *huge, const int64_t *step_dev) {
  int root = blockIdx.x, r = blockIdx.y;
  if ((mask && !mask[r]) || !mc[r]) return;
  const int64_t step = step_dev ? *step_dev : step0;   // graph replays read the step from the device
  size_t gr = (size_t)r * P.N + root;
  if (!has_exc[gr]) return;
  __shared__ int ws[MCB_LDS];
  mc_one_component(P, D, r, root, seg_n, seg, ecT, seed, step, huge, ws);
}

// The components listed by k_seg_fill (exc_list: [0] count, then
// replica * N + root), a grid-stride loop of 64-thread blocks: the same
// per-component work as k_mc_wave without a block per nucleotide.
__global__ void __launch_bounds__(64) k_mc_list(Params P, Ptrs D, int *seg_n, int *seg, const int *exc_list,
                                                const double *ecT, uint64_t seed, int64_t step0, int *huge,
                                                const int64_t *step_dev) {
  const int64_t step = step_dev ? *step_dev : step0;
  const int n = exc_list[0];
  __shared__ int ws[MCB_LDS];
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    const int gr = exc_list[1 + k];
    const int r = gr / P.N, root = gr - r * P.N;
    mc_one_component(P, D, r, root, seg_n, seg, ecT, seed, step, huge, ws);
  }
}

// ---------------------------------------------------------------------
// Parity noise: 3N (or `count`) normals per replica with the polar
// Box-Muller of src/rnd_boxmuller.F90 on the replica's MT words, replicas
// in order, sharing one cached value as the reference does.
__global__ void k_noise_parity(Params P, Ptrs D, int count, int dest, const int *mask) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int r = 0; r < P.R; r++) {
    if (mask && !mask[r]) continue;
    const uint32_t *w = D.tape + D.tape_off[r];
    int64_t pos = D.tape_pos[r], lim = D.tape_off[r + 1] - D.tape_off[r];
    double *out = D.noise + (size_t)r * P.N * 3;
    for (int k = 0; k < count; k++) {
      double val;
      if (D.bm_cache[0] == 0.0) {
        D.bm_cache[0] = 1.0;
        double vx, vy, r2;
        for (;;) {
          if (pos + 4 > lim) { D.flags[1] = 1; D.tape_pos[r] = pos; return; }
          vx = 2.0 * mt_d1(w[pos], w[pos + 1]) - 1.0;
          vy = 2.0 * mt_d1(w[pos + 2], w[pos + 3]) - 1.0;
          pos += 4;
          r2 = vx * vx + vy * vy;
          if (r2 < 1.0 && r2 > 0.0) break;
        }
        double rf = sqrt(-2.0 * log(r2) / r2);
        D.bm_cache[1] = vx * rf;
        val = vy * rf;
      } else {
        D.bm_cache[0] = 0.0;
        val = D.bm_cache[1];
      }
      out[k] = val;
    }
    D.tape_pos[r] = pos;
  }
  (void)dest;
}

// ---------------------------------------------------------------------
// Owner-computes force on bead a of replica r: bonded terms recomputed
// per owner, WCA / DH over the full neighbour lists, 6-bead BP forces
// gathered through the incidence list (no atomics on forces).
template <typename T>
__device__ __forceinline__ V3<double> bead_force(const Params &P, const Ptrs &D, int r, int a) {
  const int N = P.N;
  const double *box = D.box + 3 * r;
  const double4 *x = D.pos + (size_t)r * N;
  double4 pa = x[a];
  Topo ta = D.topo[a];
  int loc = ta.loc, n = ta.len, st = ta.start;
  bool circ = topo_circ(ta);
  V3<double> F{0.0, 0.0, 0.0};
  auto at = [&](int s, int k) -> int {  // bead of term starting at local s, position k
    int l = s + k;
    if (l > n) l -= n;
    return st + l - 1;
  };
  auto valid = [&](int &s, int m) -> bool {  // term of m beads starting at local s
    if (circ) {
      if (n < m) return false;
      if (s < 1) s += n;
      return true;
    }
    return s >= 1 && s + m - 1 <= n;
  };
  // bonds (src/list_local.F90:24-58)
  for (int p = 0; p < 2; p++) {
    int s = loc - p;
    if (!valid(s, 2)) continue;
    V3<double> v = pvec(P, box, x[at(s, 0)], x[at(s, 1)]);
    V3<double> f[2];
    bond_force<double>(P, v, f);
    F = F + f[p];
  }
  for (int p = 0; p < 3; p++) {
    int s = loc - p;
    if (!valid(s, 3)) continue;
    double4 p0 = x[at(s, 0)], p1 = x[at(s, 1)], p2 = x[at(s, 2)];
    V3<double> v12 = pvec(P, box, p0, p1), v32 = pvec(P, box, p2, p1);
    V3<double> f[3];
    angle_force<double>(P, v12, v32, f);
    F = F + f[p];
  }
  if (P.dih_type) {
    for (int p = 0; p < 4; p++) {
      int s = loc - p;
      if (!valid(s, 4)) continue;
      double4 p0 = x[at(s, 0)], p1 = x[at(s, 1)], p2 = x[at(s, 2)], p3 = x[at(s, 3)];
      V3<double> vij = pvec(P, box, p0, p1), vkj = pvec(P, box, p2, p1), vkl = pvec(P, box, p2, p3);
      V3<double> f[4];
      dih_force<double>(P, vij, vkj, vkl, f);
      F = F + f[p];
    }
  }
  size_t g = (size_t)r * N + a;
  // WCA
  {
    int nw = min(D.n_wca[g], D.maxw);
    const int *nb = D.nb_wca + (size_t)r * D.maxw * N + a;
    for (int k = 0; k < nw; k++) {
      int b = nb[(size_t)k * N];
      V3<T> v = cvt<T>(pvec(P, box, pa, x[b]));
      F = F + cvt<double>(wca_force<T>(P, v));
    }
  }
  // Debye-Hueckel
  if (P.ele_on) {
    int ne = min(D.n_ele[g], D.maxnb);
    const int *nb = D.nb_ele + (size_t)r * D.maxnb * N + a;
    T coef = (T)D.ele_coef[r], rc2 = (T)(D.ele_cut[r] * D.ele_cut[r]), rl = (T)(1.0 / D.lambdaD[r]);
    V3<T> acc{0, 0, 0};
    V3<double> accd{0.0, 0.0, 0.0};
    for (int k = 0; k < ne; k++) {
      int b = nb[(size_t)k * N];
      V3<T> v = cvt<T>(pvec(P, box, pa, x[b]));
      V3<T> f = dh_force<T>(coef, rc2, rl, v);
      accd = accd + cvt<double>(f);
    }
    (void)acc;
    F = F + accd;
  }
  // base pairs
  {
    int ni = min(D.n_inc[g], D.maxinc);
    const int *inc = D.inc + (size_t)r * D.maxinc * N + a;
    for (int k = 0; k < ni; k++) {
      int e = inc[(size_t)k * N], gb = e >> 3, role = e & 7;
      if (!D.bp_status[gb]) continue;
      const double *fo = D.for_bp + (size_t)gb * 18 + 3 * role;
      F.x += fo[0]; F.y += fo[1]; F.z += fo[2];
    }
  }
  return F;
}

// GJF-2GJ update of one bead (src/job_md.F90:415-444), or the
// Ermak-McCammon Brownian step (lines 446-456) when P.bd; g = r*N + a.
// Returns |move|^2.
__device__ __forceinline__ double move_bead(const Params &P, const Ptrs &D, size_t g, int a,
                                            const V3<double> &F, double n0, double n1, double n2) {
  const double cr = D.md_cr[g], ca = D.md_a[a];
  const double4 p = D.pos[g];
  double4 mv = D.move[g];
  double dx, dy, dz;
  if (P.bd) {
    // dx = (h/gamma) F + sqrt(2 kT h/gamma) xi; velocities are not used
    dx = ca * F.x + cr * n0;
    dy = ca * F.y + cr * n1;
    dz = ca * F.z + cr * n2;
  } else {
    const double cf = D.md_cf[a], cx = D.md_cx[a];
    double4 v = D.vel[g];
    const double4 ac = D.acc[g];
    const double b0 = cr * n0, b1 = cr * n1, b2 = cr * n2;
    v.x = ca * v.x + cf * F.x + (ac.x + b0);
    v.y = ca * v.y + cf * F.y + (ac.y + b1);
    v.z = ca * v.z + cf * F.z + (ac.z + b2);
    dx = cx * v.x; dy = cx * v.y; dz = cx * v.z;
    D.vel[g] = v;
    D.acc[g] = double4{b0, b1, b2, 0.0};
  }
  D.frc[g] = double4{F.x, F.y, F.z, 0.0};
  mv.x += dx; mv.y += dy; mv.z += dz;
  D.move[g] = mv;
  // other beads still read D.pos in this pass: write the next buffer
  D.pos_out[g] = double4{p.x + dx, p.y + dy, p.z + dz, p.w};
  return dot_nc(mv.x, mv.y, mv.z, mv.x, mv.y, mv.z);
}

// Force (and optionally the GJF-2GJ update, src/job_md.F90:415-444) per
// bead; block-level top-2 of the accumulated displacement for the skin
// test (src/job_md.F90:376-386).
template <typename T, int B>
__global__ void __launch_bounds__(B) k_force_int(Params P, Ptrs D, int integrate, int parity,
                                                uint64_t seed, int64_t step, const int *mask) {
  __shared__ double s1[B], s2[B];
  int a = blockIdx.x * B + threadIdx.x, r = blockIdx.y;
  const int N = P.N;
  double d2 = 0.0;
  bool act = a < N && !(mask && !mask[r]);
  if (act) {
    size_t g = (size_t)r * N + a;
    V3<double> F = bead_force<T>(P, D, r, a);
    D.nt_cnt[g] = -P.max_bp_per_nt;
    if (!integrate) {
      D.frc[g] = double4{F.x, F.y, F.z, 0.0};
    } else {
      double n0, n1, n2;
      if (parity) {
        const double *nz = D.noise + g * 3;
        n0 = nz[0]; n1 = nz[1]; n2 = nz[2];
      } else {
        U4 o = philox(U4{(uint32_t)step, (uint32_t)(step >> 32), (uint32_t)a, 0x6e6f6973u},
                      (uint32_t)seed, (uint32_t)(seed >> 32) ^ (0x9e3779b9u * (uint32_t)(r + 1)));
        U4 o2 = philox(U4{(uint32_t)step, (uint32_t)(step >> 32), (uint32_t)a, 0x6e6f6974u},
                       (uint32_t)seed, (uint32_t)(seed >> 32) ^ (0x9e3779b9u * (uint32_t)(r + 1)));
        double u1 = u01(o.x, o.y), u2 = u01(o.z, o.w), u3 = u01(o2.x, o2.y), u4 = u01(o2.z, o2.w);
        double rr = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
        n0 = rr * cos(6.283185307179586 * u2);
        n1 = rr * sin(6.283185307179586 * u2);
        n2 = r2 * cos(6.283185307179586 * u4);
      }
      d2 = move_bead(P, D, g, a, F, n0, n1, n2);
    }
  }
  if (!integrate) return;
  // block top-2 of d2
  s1[threadIdx.x] = d2;
  s2[threadIdx.x] = 0.0;
  __syncthreads();
  for (int off = B / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      double a1 = s1[threadIdx.x], a2 = s2[threadIdx.x], b1 = s1[threadIdx.x + off], b2 = s2[threadIdx.x + off];
      double m1 = a1 > b1 ? a1 : b1;
      double m2 = a1 > b1 ? (a2 > b1 ? a2 : b1) : (a1 > b2 ? a1 : b2);
      s1[threadIdx.x] = m1;
      s2[threadIdx.x] = m2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    D.blk_top2[((size_t)r * gridDim.x + blockIdx.x) * 2] = s1[0];
    D.blk_top2[((size_t)r * gridDim.x + blockIdx.x) * 2 + 1] = s2[0];
  }
}

// Lane-split owner-computes: LPB lanes share bead a; lane l takes items
// l, l+LPB, ... of the bead's work list [bonded terms | WCA neighbours |
// DH neighbours | BP incidences]; partial forces are reduced with lane
// shuffles (FP64).
// Bonded terms of bead a (owner-computes): lane l of an LPB-lane group
// takes terms l, l+LPB, ... of [2 bonds | 3 angles | 4 dihedrals]
// (src/list_local.F90:24-104 enumerates the same terms).  FP64 always.
__device__ __forceinline__ V3<double> bonded_part(const Params &P, const Ptrs &D, int r, int a,
                                                  int l, int LPB, const double4 *xr = nullptr) {
  const int N = P.N;
  const double *box = D.box + 3 * r;
  const double4 *x = xr ? xr : D.pos + (size_t)r * N;
  const Topo ta = D.topo[a];
  const int loc = ta.loc, n = ta.len, st = ta.start;
  const bool circ = topo_circ(ta);
  const int nbond = P.dih_type ? 9 : 5;
  V3<double> F{0.0, 0.0, 0.0};
  for (int it = l; it < nbond; it += LPB) {
    const int m = it < 2 ? 2 : (it < 5 ? 3 : 4);
    const int p = it < 2 ? it : (it < 5 ? it - 2 : it - 5);
    int s0 = loc - p;
    bool ok;
    if (circ) { ok = n >= m; if (s0 < 1) s0 += n; }
    else ok = s0 >= 1 && s0 + m - 1 <= n;
    if (!ok) continue;
    auto bead = [&](int k) { int q = s0 + k; if (q > n) q -= n; return st + q - 1; };
    if (m == 2) {
      F = F + bond_force_p(P, pvec(P, box, x[bead(0)], x[bead(1)]), p);
    } else if (m == 3) {
      const double4 p1 = x[bead(1)];
      F = F + angle_force_p(P, pvec(P, box, x[bead(0)], p1), pvec(P, box, x[bead(2)], p1), p);
    } else {
      const double4 p1 = x[bead(1)], p2 = x[bead(2)];
      F = F + dih_force_p(P, pvec(P, box, x[bead(0)], p1), pvec(P, box, p2, p1),
                          pvec(P, box, p2, x[bead(3)]), p);
    }
  }
  return F;
}

// Pair terms of bead a: WCA and DH over the full neighbour lists in T,
// 6-bead BP forces gathered through the incidence list; FP64 sums.
// Pair forces on bead a, lane l of LPB: WCA and Debye-Hueckel neighbours
// (pair_nb), then the formed base pairs it takes part in (pair_bp), both
// accumulated onto F in that order.  pair_part is the two in sequence.
template <typename T>
__device__ __forceinline__ V3<double> pair_nb(const Params &P, const Ptrs &D, int r, int a, int l, int LPB,
                                              const double4 *xr, V3<double> F) {
  const int N = P.N;
  const double *box = D.box + 3 * r;
  const double4 *x = xr ? xr : D.pos + (size_t)r * N;
  const double4 pa = x[a];
  const size_t g = (size_t)r * N + a;
  const int nw = min(D.n_wca[g], D.maxw);
  const int ne = P.ele_on ? min(D.n_ele[g], D.maxnb) : 0;
  const T coef = (T)D.ele_coef[r], rc2 = (T)(D.ele_cut[r] * D.ele_cut[r]), rl = (T)(1.0 / D.lambdaD[r]);
  // neighbours in batches of PU: their index and position loads are all
  // issued before the arithmetic; the summation order is unchanged
  constexpr int PU = 4;
  const int *nbw = D.nb_wca + (size_t)r * D.maxw * N + a;
  for (int k0 = l; k0 < nw; k0 += PU * LPB) {
    int id[PU];
    double4 pb[PU];
#pragma unroll
    for (int u = 0; u < PU; u++) {
      const int k = k0 + u * LPB;
      id[u] = k < nw ? nbw[(size_t)k * N] : -1;
    }
#pragma unroll
    for (int u = 0; u < PU; u++) pb[u] = x[id[u] >= 0 ? id[u] : a];
#pragma unroll
    for (int u = 0; u < PU; u++)
      if (id[u] >= 0) F = F + cvt<double>(wca_force<T>(P, cvt<T>(pvec(P, box, pa, pb[u]))));
  }
  const int *nbe = D.nb_ele + (size_t)r * D.maxnb * N + a;
  for (int k0 = l; k0 < ne; k0 += PU * LPB) {
    int id[PU];
    double4 pb[PU];
#pragma unroll
    for (int u = 0; u < PU; u++) {
      const int k = k0 + u * LPB;
      id[u] = k < ne ? nbe[(size_t)k * N] : -1;
    }
#pragma unroll
    for (int u = 0; u < PU; u++) pb[u] = x[id[u] >= 0 ? id[u] : a];
#pragma unroll
    for (int u = 0; u < PU; u++)
      if (id[u] >= 0) F = F + cvt<double>(dh_force<T>(coef, rc2, rl, cvt<T>(pvec(P, box, pa, pb[u]))));
  }
  return F;
}

__device__ __forceinline__ V3<double> pair_bp(const Params &P, const Ptrs &D, int r, int a, int l, int LPB,
                                              V3<double> F) {
  const int N = P.N;
  const size_t g = (size_t)r * N + a;
  const int ni = min(D.n_inc[g], D.maxinc);
  const int *inc = D.inc + (size_t)r * D.maxinc * N + a;
  for (int k = l; k < ni; k += LPB) {
    int e = inc[(size_t)k * N], gb = e >> 3, role = e & 7;
    if (D.bp_status[gb]) {
      const double *fo = D.for_bp + (size_t)gb * 18 + 3 * role;
      F.x += fo[0]; F.y += fo[1]; F.z += fo[2];
    }
  }
  return F;
}

template <typename T>
__device__ __forceinline__ V3<double> pair_part(const Params &P, const Ptrs &D, int r, int a, int l,
                                                int LPB, const double4 *xr = nullptr) {
  return pair_bp(P, D, r, a, l, LPB, pair_nb<T>(P, D, r, a, l, LPB, xr, V3<double>{0.0, 0.0, 0.0}));
}

template <typename T>
__device__ __forceinline__ V3<double> bead_force_part(const Params &P, const Ptrs &D, int r, int a,
                                                      int l, int LPB) {
  return bonded_part(P, D, r, a, l, LPB) + pair_part<T>(P, D, r, a, l, LPB);
}

// One bead's update with its noise draws (the injected MT words in the
// parity form, else the counter RNG); r indexes the per-replica arrays,
// rk keys the random stream.
__device__ __forceinline__ double gjf_update(const Params &P, const Ptrs &D, int r, int rk, int a,
                                             const V3<double> &F, int parity, uint64_t seed,
                                             int64_t step) {
  size_t g = (size_t)r * P.N + a;
  double n0, n1, n2;
  if (parity) {
    const double *nz = D.noise + g * 3;
    n0 = nz[0]; n1 = nz[1]; n2 = nz[2];
  } else {
    uint32_t k1 = (uint32_t)(seed >> 32) ^ (0x9e3779b9u * (uint32_t)(rk + 1));
    U4 o = philox(U4{(uint32_t)step, (uint32_t)(step >> 32), (uint32_t)a, 0x6e6f6973u}, (uint32_t)seed, k1);
    U4 o2 = philox(U4{(uint32_t)step, (uint32_t)(step >> 32), (uint32_t)a, 0x6e6f6974u}, (uint32_t)seed, k1);
    double u1 = u01(o.x, o.y), u2 = u01(o.z, o.w), u3 = u01(o2.x, o2.y), u4 = u01(o2.z, o2.w);
    double rr = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    n0 = rr * cos(6.283185307179586 * u2);
    n1 = rr * sin(6.283185307179586 * u2);
    n2 = r2 * cos(6.283185307179586 * u4);
  }
  return move_bead(P, D, g, a, F, n0, n1, n2);
}

template <int LPB>
__device__ __forceinline__ V3<double> group_sum(V3<double> F) {
#pragma unroll
  for (int off = LPB / 2; off > 0; off >>= 1) {
    F.x += __shfl_xor(F.x, off, LPB);
    F.y += __shfl_xor(F.y, off, LPB);
    F.z += __shfl_xor(F.z, off, LPB);
  }
  return F;
}

// bonded forces per bead into D.fbond (FP64)
template <int LPB, int B>
__global__ void __launch_bounds__(B) k_bonded(Params P, Ptrs D, const int *mask) {
  const int grp = threadIdx.x / LPB, l = threadIdx.x % LPB;
  const int a = blockIdx.x * (B / LPB) + grp, r = blockIdx.y;
  const bool act = a < P.N && !(mask && !mask[r]);
  V3<double> F{0.0, 0.0, 0.0};
  if (act) F = bonded_part(P, D, r, a, l, LPB);
  F = group_sum<LPB>(F);
  if (act && l == 0) D.fbond[(size_t)r * P.N + a] = double4{F.x, F.y, F.z, 0.0};
}

// top two of (a1 >= a2) and (b1 >= b2): the two largest values of the four
__device__ __forceinline__ void top2_merge(double &m1, double &m2, double b1, double b2) {
  const double n1 = m1 > b1 ? m1 : b1;
  const double n2 = m1 > b1 ? (m2 > b1 ? m2 : b1) : (m1 > b2 ? m1 : b2);
  m1 = n1;
  m2 = n2;
}

// Steps replayed from a captured graph (the per-phase form without a host
// round trip per step): the step each launch works on, when it stops, and
// the rebuild it stopped for.
struct GStep {
  int64_t *istep;        // [1] last completed step
  int64_t *istep_cur;    // [1] step of the current launch sequence
  const int64_t *stop;   // [1] stop once this step is completed
  int *halt;             // [1] 1: the next steps do nothing (a rebuild or the stop is due)
  int *active;           // [R] mask of the step's kernels (0 after a halt)
  int *force_mc;         // [R] limit MC on the next step (after a rebuild)
  int *mc;               // [R] limit MC on the current step
};

// the skin decision of replica r from its nblk block maxima, reduced by the
// calling workgroup (B threads, s1/s2 of B doubles); with a graph step
// also the halt and the step's commit.  Called by every thread.
template <int B>
__device__ __forceinline__ void skin_decide(const Params &P, const Ptrs &D, int r, int nblk, const GStep &G,
                                            double *s1, double *s2) {
  const int t = threadIdx.x;
  double m1 = 0.0, m2 = 0.0;
  for (int b = t; b < nblk; b += B)
    top2_merge(m1, m2, D.blk_top2[((size_t)r * nblk + b) * 2], D.blk_top2[((size_t)r * nblk + b) * 2 + 1]);
  s1[t] = m1;
  s2[t] = m2;
  __syncthreads();
  for (int off = B / 2; off > 0; off >>= 1) {
    if (t < off) {
      double a1 = s1[t], a2 = s2[t];
      top2_merge(a1, a2, s1[t + off], s2[t + off]);
      s1[t] = a1;
      s2[t] = a2;
    }
    __syncthreads();
  }
  if (t != 0) return;
  const int f = (sqrt(s1[0]) + sqrt(s2[0]) > P.nl_margin) ? 1 : 0;
  D.rebuild_flag[r] = f;
  if (G.halt) {
    if (f) atomicOr(G.halt, 1);
    if (r == 0) {
      *G.istep = *G.istep_cur;
      if (*G.istep_cur >= *G.stop) atomicOr(G.halt, 1);
    }
  }
}

// Workgroups are dispatched to the eight XCDs in turn (linear ID mod 8),
// each XCD with its own L2.  xcd_tile maps a linear workgroup ID to a tile
// so that XCD x takes the x-th contiguous slice of the n tiles: the
// position and list gathers of one XCD's workgroups then fall mostly on
// beads its own L2 already holds.  One-to-one for any n.
__device__ __forceinline__ int xcd_tile(int id, int n) {
  const int q = n >> 3, rem = n & 7, x = id & 7;
  return x * q + min(x, rem) + (id >> 3);
}

template <typename T, int LPB, int B>
__global__ void __launch_bounds__(B) k_force_lanes(Params P, Ptrs D, int integrate, int parity,
                                                  uint64_t seed, int64_t step0, const int *mask,
                                                  const int64_t *step_dev, int bonded_inline) {
  __shared__ double s1[B / LPB], s2[B / LPB];
  const int64_t step = step_dev ? *step_dev : step0;   // graph replays read the step from the device
  const int grp = threadIdx.x / LPB, l = threadIdx.x % LPB;
  // tile (bx, r) of this workgroup: grid x = tiles of a replica, y = replicas
  const int tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int bx = tile % gridDim.x, r = tile / gridDim.x;
  const int a = bx * (B / LPB) + grp;
  const int N = P.N;
  const bool act = a < N && !(mask && !mask[r]);
  // bonded forces here (bonded_inline) or from k_bonded's D.fbond (which
  // external terms add to): the same lanes and sums either way
  V3<double> Fb{0.0, 0.0, 0.0};
  if (bonded_inline) {
    if (act) Fb = bonded_part(P, D, r, a, l, LPB);
    Fb = group_sum<LPB>(Fb);
  }
  V3<double> F{0.0, 0.0, 0.0};
  if (act) F = pair_part<T>(P, D, r, a, l, LPB);
  F = group_sum<LPB>(F);
  double d2 = 0.0;
  if (act && l == 0) {
    size_t g = (size_t)r * N + a;
    const double4 fb = bonded_inline ? double4{Fb.x, Fb.y, Fb.z, 0.0} : D.fbond[g];
    F.x += fb.x; F.y += fb.y; F.z += fb.z;
    D.nt_cnt[g] = -P.max_bp_per_nt;
    if (!integrate) D.frc[g] = double4{F.x, F.y, F.z, 0.0};
    else d2 = gjf_update(P, D, r, r, a, F, parity, seed, step);
  }
  if (!integrate) return;
  if (l == 0) { s1[grp] = d2; s2[grp] = 0.0; }
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
    D.blk_top2[((size_t)r * gridDim.x + bx) * 2] = s1[0];
    D.blk_top2[((size_t)r * gridDim.x + bx) * 2 + 1] = s2[0];
  }
}

// skin test per replica (src/job_md.F90:376-395): sqrt(d2max) +
// sqrt(d2max_2nd) > nl_margin over the per-block top-2 displacements,
// reduced by a workgroup (the top two are the same values in any order).
// With a graph step: a replica needing a rebuild halts the steps that
// follow, and the step is committed.
template <int B>
__global__ void __launch_bounds__(B) k_skin(Params P, Ptrs D, int nblk, GStep G) {
  __shared__ double s1[B], s2[B];
  const int r = blockIdx.x;
  if (G.halt && !G.active[r]) return;
  skin_decide<B>(P, D, r, nblk, G, s1, s2);
}

// start of a replayed step: nothing if halted; else the step number and
// the limit-MC flags (every nstep_MC steps and after a rebuild,
// src/job_md.F90:364-394)
__global__ void k_gstep_begin(Params P, GStep G) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P.R) return;
  const bool halted = *G.halt != 0;
  G.active[r] = halted ? 0 : 1;
  if (halted) return;
  const int64_t s = *G.istep + 1;
  if (r == 0) *G.istep_cur = s;
  G.mc[r] = P.nstep_bp_mc > 0 && ((s % P.nstep_bp_mc) == 0 || G.force_mc[r]) ? 1 : 0;
  G.force_mc[r] = 0;
}

// start of a replayed step and its resets in one pass: the step number,
// limit-MC flags and mask (as k_gstep_begin), then for an active step the
// formed-pair counts and, for the performance-form MC, union-find parents,
// buckets and the list of components with excess (as k_mc_reset)
__global__ void k_gstep_reset(Params P, Ptrs D, GStep G, int *seg_n, int *has_exc, int *exc_list, int mc_reset) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  const bool halted = *G.halt != 0;
  if (i == 0) {
    G.active[r] = halted ? 0 : 1;
    if (!halted) {
      const int64_t s = *G.istep + 1;
      if (r == 0) *G.istep_cur = s;
      G.mc[r] = P.nstep_bp_mc > 0 && ((s % P.nstep_bp_mc) == 0 || G.force_mc[r]) ? 1 : 0;
      G.force_mc[r] = 0;
    }
    if (r == 0 && exc_list) exc_list[0] = 0;
  }
  if (halted || i >= P.N) return;
  const size_t g = (size_t)r * P.N + i;
  D.nt_cnt[g] = -P.max_bp_per_nt;
  if (mc_reset) {
    D.comp[g] = i;
    seg_n[g] = 0;
    has_exc[g] = 0;
  }
}

// bp_status_MC = bp_status for the replicas that did the MC this step
// (src/job_md.F90:399-401)
__global__ void k_mc_keep(Params P, Ptrs D, int nbp_total, const int *mc, const int *mask) {
  const int gb = blockIdx.x * blockDim.x + threadIdx.x;
  if (gb >= bp_count(P, D, nbp_total)) return;
  const int r = D.bp_mp[(size_t)gb * 8 + 7];
  if (mask[r] && mc[r]) D.bp_status_mc[gb] = D.bp_status[gb];
}

__global__ void k_clear_move(Params P, Ptrs D, const int *mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (i >= P.N || (mask && !mask[r])) return;
  D.move[(size_t)r * P.N + i] = double4{0, 0, 0, 0};
}

// velocities / initial noise terms (src/job_md.F90:157-176); parity uses
// the noise buffer filled twice (velocities then accelerations).
__global__ void k_init_vel(Params P, Ptrs D, int which, int parity, uint64_t seed,
                           const int *mask) {
  int a = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (a >= P.N || (mask && !mask[r])) return;
  size_t g = (size_t)r * P.N + a;
  double n0, n1, n2;
  if (parity) {
    n0 = D.noise[3 * g]; n1 = D.noise[3 * g + 1]; n2 = D.noise[3 * g + 2];
  } else {
    U4 o = philox(U4{0xffffffffu, (uint32_t)which, (uint32_t)a, 0x696e6974u}, (uint32_t)seed,
                  (uint32_t)(seed >> 32) ^ (0x9e3779b9u * (uint32_t)(r + 1)));
    U4 o2 = philox(U4{0xfffffffeu, (uint32_t)which, (uint32_t)a, 0x696e6974u}, (uint32_t)seed,
                   (uint32_t)(seed >> 32) ^ (0x9e3779b9u * (uint32_t)(r + 1)));
    double rr = sqrt(-2.0 * log(u01(o.x, o.y))), r2 = sqrt(-2.0 * log(u01(o2.x, o2.y)));
    n0 = rr * cos(6.283185307179586 * u01(o.z, o.w));
    n1 = rr * sin(6.283185307179586 * u01(o.z, o.w));
    n2 = r2 * cos(6.283185307179586 * u01(o2.z, o2.w));
  }
  if (which == 0) {
    const double BOLTZ = 1.380649e-23 * ((1.0 / 4184.0) * 6.02214076e23);
    double c = sqrt(D.tK[r] * BOLTZ / D.mass[a]);
    D.vel[g] = double4{c * n0, c * n1, c * n2, 0.0};
  } else {
    double cr = D.md_cr[g];
    D.acc[g] = double4{cr * n0, cr * n1, cr * n2, 0.0};
  }
}

// rescale velocities of replica r by f (replica exchange,
// src/replica_exchange.F90:175-213)
__global__ void k_scale_vel(Params P, Ptrs D, int r, double f, int divide) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= P.N) return;
  double4 &v = D.vel[(size_t)r * P.N + a];
  if (divide) { v.x = v.x / f; v.y = v.y / f; v.z = v.z / f; }
  else { v.x = f * v.x; v.y = f * v.y; v.z = f * v.z; }
}

// ---------------------------------------------------------------------
// Energy terms of bead a of replica r into e[1..6] (src/energy_sumup.F90
// and the terms): 1 bond, 2 angle, 3 dihedral, 4 BP, 5 WCA, 6 DH.  Each
// term is counted once: on the first bead of a bonded term, the lower bead
// of a pair, the owner (role 0) of a base-pair candidate.  coef / rlam: the
// DH coefficient and inverse Debye length to use; bp_masked: sum only the
// pairs that survived the limit MC (src/energy_bp_limit_triplet.F90:195),
// else all formed ones (src/energy_bp_triplet.F90:88).
__device__ __forceinline__ void bead_energy(const Params &P, const Ptrs &D, int r, int a, double coef,
                                            double rc, int bp_masked, const double4 *x, double e[7]) {
  const int N = P.N;
  const double *box = D.box + 3 * r;
  Topo ta = D.topo[a];
  int loc = ta.loc, n = ta.len, st = ta.start;
  bool circ = topo_circ(ta);
  auto at = [&](int k) { int l = loc + k; if (l > n) l -= n; return st + l - 1; };
  // bond starting at this bead
  if (circ || loc + 1 <= n) {
    V3<double> v = pvec(P, box, x[at(0)], x[at(1)]);
    double d = sqrt(dot(v, v));
    e[1] += 0.5 * P.bond_k * ((d - P.bond_r0) * (d - P.bond_r0));
  }
  if ((circ && n >= 3) || loc + 2 <= n) {
    V3<double> v12 = pvec(P, box, x[at(0)], x[at(1)]), v32 = pvec(P, box, x[at(2)], x[at(1)]);
    double co = dot(v32, v12) / sqrt(dot(v12, v12) * dot(v32, v32));
    if (P.angl_reb) {
      e[2] += 0.5 * P.angl_k * ((co - P.angl_co0) * (co - P.angl_co0)) / (1.0 - co * co);
    } else {
      if (co > 1.0) co = 1.0; else if (co < -1.0) co = -1.0;
      double t = acos(co);
      e[2] += 0.5 * P.angl_k * ((t - P.angl_t0) * (t - P.angl_t0));
    }
  }
  if (P.dih_type && ((circ && n >= 4) || loc + 3 <= n)) {
    V3<double> v12 = pvec(P, box, x[at(1)], x[at(0)]), v23 = pvec(P, box, x[at(2)], x[at(1)]),
               v34 = pvec(P, box, x[at(3)], x[at(2)]);
    V3<double> c1 = cross(v12, v23), c2 = cross(v23, v34);
    double dih = atan2(sqrt(dot(v23, v23)) * dot(v12, c2), dot(c1, c2));
    if (P.dih_type == 2) e[3] -= P.dih_k * exp(-0.5 * P.dih_w * ((dih - P.dih_p0) * (dih - P.dih_p0)));
    else e[3] += P.dih_k * (1.0 + cos(dih + P.dih_p0));
  }
  size_t g = (size_t)r * N + a;
  {
    int ni = min(D.n_inc[g], D.maxinc);
    const int *inc = D.inc + (size_t)r * D.maxinc * N + a;
    for (int k = 0; k < ni; k++) {
      int q = inc[(size_t)k * N];
      if ((q & 7) == 0 && (D.bp_status[q >> 3] || !bp_masked)) e[4] += D.ene_bp[q >> 3];
    }
  }
  {
    int nw = min(D.n_wca[g], D.maxw);
    const int *nb = D.nb_wca + (size_t)r * D.maxw * N + a;
    for (int k = 0; k < nw; k++) {
      int b = nb[(size_t)k * N];
      if (b < a) continue;
      V3<double> v = pvec(P, box, x[a], x[b]);
      double d = sqrt(dot(v, v));
      if (d >= P.wca_sigma) continue;
      d = P.wca_sigma / d;
      double d6 = d * d * d * d * d * d;
      e[5] += P.wca_eps * (d6 * d6 - 2 * d6 + 1.0);
    }
  }
  if (P.ele_on) {
    int ne = min(D.n_ele[g], D.maxnb);
    const int *nb = D.nb_ele + (size_t)r * D.maxnb * N + a;
    for (int k = 0; k < ne; k++) {
      int b = nb[(size_t)k * N];
      if (b < a) continue;
      V3<double> v = pvec(P, box, x[a], x[b]);
      double dist = sqrt(dot(v, v));
      if (dist > D.ele_cut[r]) continue;
      e[6] += coef / dist * exp(-dist * rc);
    }
  }
}

// Energies (src/energy_sumup.F90): per-block partial sums of bead_energy,
// reduced in a fixed order.  r_only >= 0: one replica, grid.y = 1, its DH
// coefficient and Debye length by value (coef1, lam1) and its partial
// sums in slot 0 of `partial`.
template <int B>
__global__ void __launch_bounds__(B) k_energy(Params P, Ptrs D, const double *ecT,
                                              const double *eCoef, const double *eLam,
                                              double *partial, const int *mask, int bp_masked,
                                              int r_only, double coef1, double lam1) {
  __shared__ double s[7][B];
  int a = blockIdx.x * B + threadIdx.x, r = r_only >= 0 ? r_only : blockIdx.y;
  const int slot = r_only >= 0 ? 0 : r;
  const int N = P.N;
  double e[7] = {0, 0, 0, 0, 0, 0, 0};
  if (a < N && !(mask && !mask[r])) {
    double coef = r_only >= 0 ? coef1 : (eCoef ? eCoef[r] : D.ele_coef[r]);
    double rc = 1.0 / (r_only >= 0 ? lam1 : (eLam ? eLam[r] : D.lambdaD[r]));
    bead_energy(P, D, r, a, coef, rc, bp_masked, D.pos + (size_t)r * N, e);
  }
  (void)ecT;
  for (int k = 1; k < 7; k++) s[k][threadIdx.x] = e[k];
  __syncthreads();
  for (int off = B / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 1; k < 7; k++) s[k][threadIdx.x] += s[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 1; k < 7; k++) partial[((size_t)slot * gridDim.x + blockIdx.x) * 8 + k] = s[k][0];
}
// energy_kinetic (src/energy_kinetic.F90:16-29): 0.5 sum m |v|^2 of
// replica blockIdx.x, one workgroup per replica, FP64 tree sum.
template <int B>
__global__ void __launch_bounds__(B) k_kinetic(Params P, Ptrs D, double *out) {
  __shared__ double s[B];
  const int r = blockIdx.x, N = P.N;
  double e = 0.0;
  for (int a = threadIdx.x; a < N; a += B) {
    const double4 v = D.vel[(size_t)r * N + a];
    e += D.mass[a] * (v.x * v.x + v.y * v.y + v.z * v.z);
  }
  s[threadIdx.x] = e;
  __syncthreads();
  for (int off = B / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[r] = 0.5 * s[0];
}

// ext: stage, tweezers, Rg and restraint energies [4*R], or null.
// r_only >= 0: that replica alone, partial sums and output in slot 0.
__global__ void k_energy_reduce(const double *partial, int nblk, double *out, const double *ext, int r_only) {
  const int r = r_only >= 0 ? r_only : blockIdx.x, slot = r_only >= 0 ? 0 : r;
  if (threadIdx.x != 0) return;
  double e[11] = {0};
  for (int b = 0; b < nblk; b++)
    for (int k = 1; k < 7; k++) e[k] += partial[((size_t)slot * nblk + b) * 8 + k];
  if (ext)
    for (int k = 0; k < 4; k++) e[7 + k] = ext[4 * r + k];
  double t = 0.0;
  for (int k = 1; k <= 10; k++) t += e[k];
  e[0] = t;
  for (int k = 0; k < 11; k++) out[(size_t)slot * 16 + k] = e[k];
}

}  // namespace sis
