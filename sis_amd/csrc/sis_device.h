// Device-side data layout and per-interaction math of the SIS MD step for
// gfx950.  Included by sis_gpu.hip only.
//
// Layout (one context = R replicas of the same N-bead topology):
//   bead state   double4 pos/vel/acc/move[R*N]  (x, y, z, unused): one 32-byte
//                sector per bead for neighbour gathers, coalesced streams for
//                the integrator
//   full lists   nb_ele[R][maxnb][N], nb_wca[R][maxw][N]  (slot-major: lane i
//                reads slot k of bead i next to lane i+1 -> coalesced)
//   BP list      bp_mp[R][maxbp][8] (7 reference indices + pad), reference
//                order (src/neighbor_list.F90:99-114)
//   BP incidence inc[R][maxinc][N] = 8*b + role, for owner-computes of the
//                6-bead forces (no atomics on forces)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef __HIP_DEVICE_COMPILE__
#include <cmath>
#endif

namespace sis {

enum { SQ_A = 0, SQ_U = 1, SQ_G = 2, SQ_C = 3, SQ_D = 4 };

struct BpPara {
  double cutoff_ddist, U0, bond_k, bond_r;
  double angl_k1, angl_k2, angl_k3, angl_k4;
  double angl_theta1, angl_theta2, angl_theta3, angl_theta4;
  double dihd_k1, dihd_k2, dihd_phi1, dihd_phi2;
};

// Scalars shared by all replicas of a context (passed by value).
struct Params {
  double bond_k, bond_r0, angl_k, angl_t0, angl_co0;
  double dih_k, dih_w, dih_p0;
  double wca_sigma, wca_eps, wca_sig2, wca_coef;
  double bp_cutoff_energy, bp_cutoff_dist, nl_margin;
  double wca_nl_cut2, bp_nl_cut2;
  double mc_kT;   // > 0: kT of the limit-MC acceptance when it is not the replica's (annealing)
  double ele_cut_inp;   // [Electrostatic] cutoff as read (A, or Debye lengths with cutoff_type 2)
  int angl_reb, dih_type, ele_on, ele_excl_bond, ele_excl_angl;
  int bp_model, bp_min_loop, max_bp_per_nt, nstep_bp_mc, pbc;
  int N, R, fix_h4, nchains, maxlen;
  int bd;  // 1: Ermak-McCammon Brownian update in place of GJF-2GJ
  int ele_cut_type;   // 1: fixed cutoff; 2: cutoff = ele_cut_inp * lambdaD
  // 0: pair rules of the MD job's neighbour list (src/neighbor_list.F90);
  // 1: those of the cutoff-free lists of the DCD and CHECK_FORCE jobs
  // (src/list_exv.F90, src/list_ele.F90, src/list_bp.F90)
  int list_rule;
};

// 1/kT of the limit-MC acceptance.  The force's MC takes var_state's kT,
// which an annealed run never updates (src/force_bp_limit_triplet.F90:61-67);
// energy_sumup's MC takes the temperature it is given
// (src/energy_bp_limit_triplet.F90:37-39).
__device__ __forceinline__ double mc_beta(const Params &P, double tK, bool force_side) {
  const double BOLTZ = 1.380649e-23 * ((1.0 / 4184.0) * 6.02214076e23);
  return (force_side && P.mc_kT > 0.0) ? 1.0 / P.mc_kT : 1.0 / (BOLTZ * tK);
}

// Per-bead topology word: chain start (replica-local 0-based), chain
// length, local index (1-based), packed flags.
struct Topo {
  int start, len, loc, flags;  // flags: bit0 circular, bits4-7 seq, bit8 charged
};
__host__ __device__ inline int topo_seq(const Topo &t) { return (t.flags >> 4) & 15; }
__host__ __device__ inline bool topo_circ(const Topo &t) { return t.flags & 1; }
__host__ __device__ inline bool topo_charged(const Topo &t) { return (t.flags >> 8) & 1; }

// Device pointers of a context.
struct Ptrs {
  const Topo *topo;          // [N]
  const int *chain_id;       // [N] chain index of each bead
  const int *bp3_ctx;        // [N] packed (seq[i-1],seq[i],seq[i+1]) context for eligibility
  const int *bp3_hash;       // [4097] triplet hash -> table index (0 = none)
  const double *bp3_dH, *bp3_dS;  // [469]
  const BpPara *bpp;         // [7] base-pair types 1..6 = GC AU GU CG UA UG (device memory:
                             // indexed per candidate, so not part of the kernel arguments)
  const double *md_a, *md_cf, *md_cx, *mass;  // [N]
  double *md_cr;             // [R*N]
  double *tK, *ele_coef, *lambdaD, *ele_cut, *box;  // [R], box [R*3]
  double4 *pos, *vel, *acc, *move, *frc;  // [R*N]
  double4 *pos_out;          // [R*N] positions after the update (double buffer)
  double4 *fbond;            // [R*N] bonded force of each bead (computed before the pair pass)
  // lists
  int maxnb, maxw, maxbp, maxinc, maxbpb;
  int *nb_ele, *n_ele, *nb_wca, *n_wca;   // [R*maxnb*N], [R*N] ...
  int *bp_tmp, *n_bp_tmp;    // per-bead BP partners found by the search [R*N*maxbpb]
  int *bp_off;               // [R*(N+1)] exclusive scan of n_bp_tmp
  int *bp_mp;                // [R*maxbp*8]
  unsigned long long *bp_key; // [maxbp] sort key of each candidate
  int *bp_val;               // [maxbp]
  double *bp_coef;           // [R*maxbp*2]
  int *n_bp;                 // [R]
  int *inc, *n_inc;          // [R*maxinc*N], [R*N]
  uint8_t *bp_status, *bp_status_mc;  // [R*maxbp]
  double *ene_bp;            // [R*maxbp]
  double *for_bp;            // [R*maxbp*18]
  int *nt_cnt;               // [R*N] formed BPs per nt minus max_bp_per_nt
  int *comp;                 // [R*N] union-find parent
  int *flags;                // [8] 0: capacity overflow, 1: error
  // cell list
  int *cell_of, *cell_cnt, *cell_start, *cell_beads;  // cell_start [R*(ncell+1)]
  int ncell_x, ncell_y, ncell_z, ncell;
  // skin test
  double *blk_top2;          // [R*nblk*2]
  int *rebuild_flag;         // [R]
  // parity RNG
  const uint32_t *tape;      // [R] tapes concatenated
  const int64_t *tape_off;   // [R+1]
  int64_t *tape_pos;         // [R]
  double *noise;             // [R*N*3]
  double *bm_cache;          // [2] stored flag, value (process-global cache)
  // energies
  double *ene_acc;           // [R*16]
  // pairs allowed by a secondary-structure file, or null: per bead i the
  // sorted partners j > i in ss_partner[ss_off[i] .. ss_off[i+1])
  const int *ss_off, *ss_partner;
};

// Debye-Hueckel cutoff a list rebuild of replica r takes.  With cutoff_type 2
// it follows the replica's current Debye length (src/neighbor_list.F90:80-82);
// the forces and energies between rebuilds keep the value of the last
// rebuild (ele_cut), also after an exchange has changed lambdaD
// (set_ele leaves ele_cutoff alone, src/job_md.F90:556).
__device__ __forceinline__ double rebuild_ele_cut(const Params &P, const Ptrs &D, int r) {
#pragma clang fp contract(off)
  return P.ele_cut_type == 2 ? P.ele_cut_inp * D.lambdaD[r] : D.ele_cut[r];
}

// ---------------------------------------------------------------------
// minimum image (src/pbc.F90:140-165): one conditional shift per axis.
// Contraction is disabled here so that d2 matches the reference exactly.
__device__ __forceinline__ void pbc_vec(const double *box, int pbc, double ax, double ay,
                                        double az, double bx, double by, double bz,
                                        double &vx, double &vy, double &vz) {
#pragma clang fp contract(off)
  vx = ax - bx; vy = ay - by; vz = az - bz;
  if (pbc) {
    double hx = 0.5 * box[0], hy = 0.5 * box[1], hz = 0.5 * box[2];
    if (vx > hx) vx = vx - box[0]; else if (vx < -hx) vx = vx + box[0];
    if (vy > hy) vy = vy - box[1]; else if (vy < -hy) vy = vy + box[1];
    if (vz > hz) vz = vz - box[2]; else if (vz < -hz) vz = vz + box[2];
  }
}
__device__ __forceinline__ double dot_nc(double ax, double ay, double az, double bx, double by,
                                         double bz) {
#pragma clang fp contract(off)
  return ax * bx + ay * by + az * bz;
}

template <typename T>
struct V3 {
  T x, y, z;
};
template <typename T>
__device__ __forceinline__ V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T>
__device__ __forceinline__ T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
__device__ __forceinline__ V3<T> cross(const V3<T> &a, const V3<T> &b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T>
__device__ __forceinline__ V3<T> operator*(T s, const V3<T> &a) { return V3<T>{s * a.x, s * a.y, s * a.z}; }
template <typename T>
__device__ __forceinline__ V3<T> operator+(const V3<T> &a, const V3<T> &b) { return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T>
__device__ __forceinline__ V3<T> operator-(const V3<T> &a, const V3<T> &b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
__device__ __forceinline__ V3<T> operator-(const V3<T> &a) { return V3<T>{-a.x, -a.y, -a.z}; }

__device__ __forceinline__ float m_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float m_exp(float x) { return __expf(x); }
__device__ __forceinline__ double m_exp(double x) { return exp(x); }
__device__ __forceinline__ float m_acos(float x) { return acosf(x); }
__device__ __forceinline__ double m_acos(double x) { return acos(x); }
__device__ __forceinline__ float m_atan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ double m_atan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float m_sin(float x) { return sinf(x); }
__device__ __forceinline__ double m_sin(double x) { return sin(x); }
__device__ __forceinline__ float m_cos(float x) { return cosf(x); }
__device__ __forceinline__ double m_cos(double x) { return cos(x); }

__device__ __forceinline__ V3<double> pvec(const Params &P, const double *box, const double4 &a,
                                           const double4 &b) {
  V3<double> v;
  pbc_vec(box, P.pbc, a.x, a.y, a.z, b.x, b.y, b.z, v.x, v.y, v.z);
  return v;
}
template <typename T, typename S>
__device__ __forceinline__ V3<T> cvt(const V3<S> &v) { return V3<T>{(T)v.x, (T)v.y, (T)v.z}; }

// ---------------------------------------------------------------------
// Bonded terms.  Each returns the force on every bead of the term, in the
// reference's expressions (src/force_bond.F90:42-50,
// src/force_angl_ReB.F90:43-59, src/force_angl.F90:42-66,
// src/force_dih_exp.F90:44-79, src/force_dih_cos.F90:28-54).
template <typename T>
__device__ __forceinline__ void bond_force(const Params &P, const V3<T> &v, V3<T> f[2]) {
  T d = m_sqrt(dot(v, v));
  T delta = d - (T)P.bond_r0;
  V3<T> fv = ((T)P.bond_k * delta / d) * v;   // v = x1 - x2
  f[0] = -fv;
  f[1] = fv;
}
template <typename T>
__device__ __forceinline__ void angle_force(const Params &P, const V3<T> &v12, const V3<T> &v32,
                                            V3<T> f[3]) {
  if (P.angl_reb) {
    T d12 = dot(v12, v12), d32 = dot(v32, v32), a1232 = m_sqrt(d12 * d32);
    T cosine = dot(v12, v32) / a1232;
    T co0 = (T)P.angl_co0;
    T den = (T)1 - cosine * cosine;
    T ff = -(T)P.angl_k * (cosine - co0) * ((T)1 - cosine * co0) / (den * den);
    V3<T> f1 = ff * ((T(1) / a1232) * v32 - (cosine / d12) * v12);
    V3<T> f3 = ff * ((T(1) / a1232) * v12 - (cosine / d32) * v32);
    f[0] = f1; f[1] = -(f1 + f3); f[2] = f3;
  } else {
    V3<T> v21 = -v12;  // x2 - x1
    T d21 = dot(v21, v21), d32 = dot(v32, v32), d2132 = dot(v32, v21);
    T cosine = -d2132 / m_sqrt(d21 * d32);
    if (cosine > (T)1) cosine = (T)1;
    if (cosine < (T)-1) cosine = (T)-1;
    T delta = m_acos(cosine) - (T)P.angl_t0;
    T t = d21 * d32 - d2132 * d2132;
    if (t <= (T)1) t = (T)1;
    T ff = (T)P.angl_k * delta / m_sqrt(t);
    V3<T> f21 = ff * ((d2132 / d21) * v21 - v32);
    V3<T> f32 = ff * ((d2132 / d32) * v32 - v21);
    f[0] = -f21; f[1] = f21 - f32; f[2] = f32;
  }
}
template <typename T>
__device__ __forceinline__ void dih_force(const Params &P, const V3<T> &vij, const V3<T> &vkj,
                                          const V3<T> &vkl, V3<T> f[4]) {
  V3<T> m = cross(vij, vkj), n = cross(vkj, vkl);
  T akj2 = dot(vkj, vkj), akj = m_sqrt(akj2);
  T dih = m_atan2(akj * dot(vij, n), dot(m, n));
  T pre;
  if (P.dih_type == 2) {
    T dd = dih - (T)P.dih_p0;
    pre = (T)(P.dih_k * P.dih_w) * dd * m_exp((T)(-0.5 * P.dih_w) * (dd * dd)) * akj;
  } else {
    pre = -(T)P.dih_k * m_sin(dih + (T)P.dih_p0) * akj;
  }
  V3<T> fi = (-pre / dot(m, m)) * m;
  V3<T> fl = (pre / dot(n, n)) * n;
  T a = dot(vij, vkj) / akj2, b = dot(vkl, vkj) / akj2;
  f[0] = fi;
  f[1] = ((T)-1 + a) * fi + (-b) * fl;
  f[2] = (-a) * fi + ((T)-1 + b) * fl;
  f[3] = fl;
}

// Force on position p of a bonded term only (owner-computes).
__device__ __forceinline__ V3<double> bond_force_p(const Params &P, const V3<double> &v, int p) {
  double d = sqrt(dot(v, v));
  double delta = d - P.bond_r0;
  V3<double> fv = (P.bond_k * delta / d) * v;
  return p == 0 ? -fv : fv;
}
__device__ __forceinline__ V3<double> angle_force_p(const Params &P, const V3<double> &v12,
                                                    const V3<double> &v32, int p) {
  V3<double> f1, f3;
  if (P.angl_reb) {
    double d12 = dot(v12, v12), d32 = dot(v32, v32), a1232 = sqrt(d12 * d32);
    double cosine = dot(v12, v32) / a1232;
    double co0 = P.angl_co0;
    double den = 1.0 - cosine * cosine;
    double ff = -P.angl_k * (cosine - co0) * (1.0 - cosine * co0) / (den * den);
    f1 = ff * ((1.0 / a1232) * v32 - (cosine / d12) * v12);
    f3 = ff * ((1.0 / a1232) * v12 - (cosine / d32) * v32);
  } else {
    V3<double> v21 = -v12;
    double d21 = dot(v21, v21), d32 = dot(v32, v32), d2132 = dot(v32, v21);
    double cosine = -d2132 / sqrt(d21 * d32);
    if (cosine > 1.0) cosine = 1.0;
    if (cosine < -1.0) cosine = -1.0;
    double delta = acos(cosine) - P.angl_t0;
    double t = d21 * d32 - d2132 * d2132;
    if (t <= 1.0) t = 1.0;
    double ff = P.angl_k * delta / sqrt(t);
    V3<double> f21 = ff * ((d2132 / d21) * v21 - v32);
    V3<double> f32 = ff * ((d2132 / d32) * v32 - v21);
    f1 = -f21;
    f3 = f32;
  }
  if (p == 0) return f1;
  if (p == 2) return f3;
  return -(f1 + f3);
}
__device__ __forceinline__ V3<double> dih_force_p(const Params &P, const V3<double> &vij,
                                                  const V3<double> &vkj, const V3<double> &vkl, int p) {
  V3<double> m = cross(vij, vkj), n = cross(vkj, vkl);
  double akj2 = dot(vkj, vkj), akj = sqrt(akj2);
  double dih = atan2(akj * dot(vij, n), dot(m, n));
  double pre;
  if (P.dih_type == 2) {
    double dd = dih - P.dih_p0;
    pre = (P.dih_k * P.dih_w) * dd * exp((-0.5 * P.dih_w) * (dd * dd)) * akj;
  } else {
    pre = -P.dih_k * sin(dih + P.dih_p0) * akj;
  }
  V3<double> fi = (-pre / dot(m, m)) * m;
  V3<double> fl = (pre / dot(n, n)) * n;
  if (p == 0) return fi;
  if (p == 3) return fl;
  double a = dot(vij, vkj) / akj2, b = dot(vkl, vkj) / akj2;
  if (p == 1) return (-1.0 + a) * fi + (-b) * fl;
  return (-a) * fi + (-1.0 + b) * fl;
}

// WCA (src/force_wca.F90:51-58): force on bead 1 of v12 = x1 - x2.
template <typename T>
__device__ __forceinline__ V3<T> wca_force(const Params &P, const V3<T> &v12) {
  T d2 = dot(v12, v12);
  if (d2 >= (T)P.wca_sig2) return V3<T>{0, 0, 0};
  // FP32: hardware reciprocal (1 ulp) instead of the correctly rounded divide
  T s;
  if constexpr (sizeof(T) == 4) s = (T)P.wca_sig2 * __builtin_amdgcn_rcpf(d2);
  else s = (T)P.wca_sig2 / d2;
  T s2 = s * s, s4 = s2 * s2, s7 = s4 * s2 * s;
  return ((T)P.wca_coef * (s7 - s4)) * v12;
}
// Debye-Hueckel (src/force_ele_DH.F90:57-68): force on bead 1 of v12.
template <typename T>
__device__ __forceinline__ V3<T> dh_force(T coef, T rc2, T rlam, const V3<T> &v12) {
  T d2 = dot(v12, v12);
  if (d2 > rc2) return V3<T>{0, 0, 0};
  // FP32: hardware reciprocal square root (1 ulp) gives 1/d and d = d2/d
  T d1, r1;
  if constexpr (sizeof(T) == 4) {
    r1 = __builtin_amdgcn_rsqf(d2);
    d1 = d2 * r1;
  } else {
    d1 = m_sqrt(d2);
    r1 = (T)1 / d1;
  }
  T dv = coef * r1 * r1 * (r1 + rlam) * m_exp(-d1 * rlam);
  return dv * v12;
}

// One BP angle term for the angle between a (= v12) and b, with
// aa = |a|^2, bb = |b|^2, ab = a.b, nab = |a||b|.  Returns d = theta -
// theta0 and Fa = pre (a - (ab/bb) b), Fb = pre (b - (ab/aa) a) with
// pre = 2 k d / (|a||b| sin theta).  FP64 follows the reference's
// expressions (src/force_bp_limit_triplet.F90:144-154).  FP32 takes theta,
// |a||b| sin theta and the two perpendicular components from c = a x b
// instead: the same quantities, but acos, the radicand and the
// subtractions lose all accuracy (or leave their domains) for nearly
// collinear beads at FP32 rounding.
template <typename T>
__device__ __forceinline__ T bp_angle(const V3<T> &a, const V3<T> &b, T aa, T bb, T ab, T nab, T theta0, T k,
                                      V3<T> &Fa, V3<T> &Fb) {
  T d;
  if constexpr (sizeof(T) == 8) {
    d = m_acos(ab / nab) - theta0;
    const T pre = (T)2 * k * d / m_sqrt(bb * aa - ab * ab);
    Fa = pre * (a - (ab / bb) * b);
    Fb = pre * (b - (ab / aa) * a);
  } else {
    const V3<T> c = cross(a, b);
    const T s = m_sqrt(dot(c, c));
    d = m_atan2(s, ab) - theta0;
    const T g = (T)2 * k * d / fmaxf(s, 1e-30f);
    Fa = (g / bb) * cross(b, c);
    Fb = (g / aa) * cross(c, a);
  }
  return d;
}

// ---------------------------------------------------------------------
// 6-bead base-pair term (src/force_bp_limit_triplet.F90:100-243): returns
// 0 if |r - r_bp| > cutoff_ddist, else u and the unscaled forces.
// The distance test is done in FP64 whatever T is, so the formed set does
// not depend on the pair precision.
template <typename T>
__device__ __forceinline__ int bp_geom(const BpPara &bp, const V3<double> &v12d,
                                       const V3<double> &v13d, const V3<double> &v15d,
                                       const V3<double> &v42d, const V3<double> &v62d, T &u_out,
                                       V3<T> fb[6]) {
  double d1212d = dot_nc(v12d.x, v12d.y, v12d.z, v12d.x, v12d.y, v12d.z);
  double a12d = m_sqrt(d1212d);
  double dd = a12d - bp.bond_r;
  if (fabs(dd) > bp.cutoff_ddist) return 0;
  V3<T> v12 = cvt<T>(v12d), v13 = cvt<T>(v13d), v15 = cvt<T>(v15d), v42 = cvt<T>(v42d),
        v62 = cvt<T>(v62d);
  T d1212 = (T)d1212d, a12 = (T)a12d, d = (T)dd;
  T u = (T)bp.bond_k * (d * d);
  V3<T> fi = ((T)2 * (T)bp.bond_k * d / a12) * v12, fk;
  fb[0] = fi;
  fb[1] = -fi;
  T d1313 = dot(v13, v13), d4242 = dot(v42, v42), d1213 = dot(v12, v13), d1242 = dot(v12, v42);
  T d1215 = dot(v12, v15), d1515 = dot(v15, v15), d1262 = dot(v12, v62), d6262 = dot(v62, v62);
  T o13 = d1213 / d1212, o42 = d1242 / d1212, o15 = d1215 / d1212, o62 = d1262 / d1212;
  V3<T> fa, fc;
  T pre;
  // angle 3-1=2
  d = bp_angle<T>(v12, v13, d1212, d1313, d1213, m_sqrt(d1313) * a12, (T)bp.angl_theta2, (T)bp.angl_k2, fa, fc);
  u = u + (T)bp.angl_k2 * (d * d);
  fi = fa;
  fk = fc;
  fb[0] = fb[0] - fi - fk; fb[1] = fb[1] + fk; fb[2] = fi;
  // angle 1=2-4
  d = bp_angle<T>(v12, v42, d1212, d4242, d1242, a12 * m_sqrt(d4242), (T)bp.angl_theta1, (T)bp.angl_k1, fa, fc);
  u = u + (T)bp.angl_k1 * (d * d);
  fi = -fc;
  fk = -fa;
  fb[0] = fb[0] + fi; fb[1] = fb[1] - fi - fk; fb[3] = fk;
  // angle 5-1=2
  d = bp_angle<T>(v12, v15, d1212, d1515, d1215, m_sqrt(d1515) * a12, (T)bp.angl_theta4, (T)bp.angl_k4, fa, fc);
  u = u + (T)bp.angl_k4 * (d * d);
  fi = fa;
  fk = fc;
  fb[0] = fb[0] - fi - fk; fb[1] = fb[1] + fk; fb[4] = fi;
  // angle 1=2-6
  d = bp_angle<T>(v12, v62, d1212, d6262, d1262, a12 * m_sqrt(d6262), (T)bp.angl_theta3, (T)bp.angl_k3, fa, fc);
  u = u + (T)bp.angl_k3 * (d * d);
  fi = -fc;
  fk = -fa;
  fb[0] = fb[0] + fi; fb[1] = fb[1] - fi - fk; fb[5] = fk;
  // dihedral 4-2=1=3
  V3<T> m = cross(v42, v12), n = cross(v12, v13), fl;
  T dih = m_atan2(dot(v42, n) * a12, dot(m, n));
  d = dih + (T)bp.dihd_phi1;
  u = u + (T)bp.dihd_k1 * ((T)1 + m_cos(d));
  pre = -(T)bp.dihd_k1 * m_sin(d) * a12;
  fi = (pre / dot(m, m)) * m;
  fl = (-pre / dot(n, n)) * n;
  fb[3] = fb[3] + fi;
  fb[1] = fb[1] + ((T)-1 + o42) * fi - o13 * fl;
  fb[0] = fb[0] + ((T)-1 + o13) * fl - o42 * fi;
  fb[2] = fb[2] + fl;
  // dihedral 6-2=1=5
  m = cross(v62, v12); n = cross(v12, v15);
  dih = m_atan2(dot(v62, n) * a12, dot(m, n));
  d = dih + (T)bp.dihd_phi2;
  u = u + (T)bp.dihd_k2 * ((T)1 + m_cos(d));
  pre = -(T)bp.dihd_k2 * m_sin(d) * a12;
  fi = (pre / dot(m, m)) * m;
  fl = (-pre / dot(n, n)) * n;
  fb[5] = fb[5] + fi;
  fb[1] = fb[1] + ((T)-1 + o62) * fi - o15 * fl;
  fb[0] = fb[0] + ((T)-1 + o15) * fl - o62 * fi;
  fb[4] = fb[4] + fl;
  u_out = u;
  return 1;
}

// ---------------------------------------------------------------------
// Random numbers.
// Parity mode: tempered MT19937 words -> genrand_double1 / genrand_double3
// (mt_stream.F90:708-729, 754-775).
__device__ __forceinline__ double mt_d1(uint32_t w1, uint32_t w2) {
  double a = (double)(w1 >> 5), b = (double)(w2 >> 6);
  return (a * 67108864.0 + b) * (1.0 / 9007199254740991.0);
}
__device__ __forceinline__ double mt_d3(uint32_t w1, uint32_t w2) {
  double a = (double)(w1 >> 6), b = (double)(w2 >> 6);
  return (a * 67108864.0 + b + 0.5) * (1.0 / 4503599627370496.0);
}

// Performance mode: Philox4x32-10 (Salmon et al., SC'11) keyed by
// (seed, replica); counter = (step, stream, index, 0).
struct U4 { uint32_t x, y, z, w; };
__host__ __device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
// uniform in (0,1) with 53 bits from two words
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  m &= ((1ull << 53) - 1);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

}  // namespace sis
