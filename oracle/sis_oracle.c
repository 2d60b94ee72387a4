/*
 * sis_oracle.c - CPU restatement of the SIS MD hot path.  TEST ORACLE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
 * bench.py; never by the product library.
 *
 * Restated from the reference Fortran (file:line in each comment).  The
 * expressions keep the reference's operation order so that, built with
 * -ffp-contract=off, results agree with the reference to the last few ulps
 * (libm is glibc on both sides).  Floating point is FP64 throughout
 * (PREC = REAL64, src/const.F90:10).
 */
#include "sis_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- physical constants (src/const_phys.F90:5-14) ---- */
#define KCAL2JOUL 4184.0
static const double PI_ = 3.141592653589793115997963468544185161590576171875;
static const double EPS0 = 8.8541878128e-12, ELE = 1.602176634e-19;
static const double BOLTZ_J = 1.380649e-23, N_AVO = 6.02214076e23;
static double JOUL2KCAL_MOL, BOLTZ_KCAL_MOL;

enum { A_ = 0, U_ = 1, G_ = 2, C_ = 3, D_ = 4 };
enum { GC = 1, AU = 2, GU = 3, CG = 4, UA = 5, UG = 6 };

typedef struct { int32_t *a; int64_t n, cap; } ivec;
typedef struct { double *a; int64_t n, cap; } dvec;

static void ipush(ivec *v, int32_t x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 1024; v->a = realloc(v->a, v->cap * sizeof(int32_t)); }
    v->a[v->n++] = x;
}
static void dpush(dvec *v, double x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 1024; v->a = realloc(v->a, v->cap * sizeof(double)); }
    v->a[v->n++] = x;
}

typedef struct {
    const uint32_t *w; int64_t n, pos;
} tape_t;

typedef struct {
    ivec wca, ele, bp;       /* wca/ele: 2 per pair; bp: 7 per candidate */
    dvec bpcoef;             /* 2 per candidate */
    uint8_t *bp_status, *bp_status_mc;
    double *ene_bp;
    int have_status_mc;
    double lambdaD, ele_coef, ele_cutoff, diele;
    double tK;               /* temperature of this replica */
    double ionic;            /* ionic strength of this replica */
    tape_t tape;
} rep_t;

struct orc_sys {
    sis_params p;
    sis_bp_para bpp[7];      /* 1..6 */
    int nmp, nchains, nrep;
    int *nmp_chain, *moltype, *seq, *chain_of, *lmp, *chain_start;
    uint8_t *has_charge;
    double *xyz, *velos, *accels, *move;  /* [nrep][nmp][3] */
    double *mass, *fric;
    double bp_cutoff_dist;
    double bp3_dH[469], bp3_dS[469];
    int bp3_hash[4097];
    rep_t *rep;
    /* rnd_boxmuller process-global cache (src/rnd_boxmuller.F90:13-15) */
    int bm_stored; double bm_value;
    /* replica exchange */
    int *rep2lab, *lab2rep; double *lab2val; int idx_pair; int *pair_tb; /* [pair_sets][nrep+2] */
    int pair_sets;           /* 2 x ndim_replica */
    double *lab2ion;         /* ionic strength of each label (REMD over REPT%ION) */
    double *lab2rg;          /* Rg-bias target of each label (REMD over REPT%RG) */
    int rg_remd;             /* the Rg targets travel with the labels */
    double *lab2dcf; double dcf_dir[3]; int dcf_remd;   /* f-REMD (REPT%TWZDCF) */
    tape_t tape_rep;
    int64_t istep;
    int flg_bp_energy;
    /* planted mistakes, for tests that show a fixture catches them (orc_plant) */
    int plant_cut_at_set_ele;
    /* counter-based random numbers in place of the MT tapes (orc_set_counter_rng) */
    int ctr_on; uint64_t ctr_seed;
    int nl_brute;            /* 1: always the O(N^2) loops (orc_set_nl_brute) */
    /* external terms (copies of the caller's arrays) */
    int x_on;
    sis_external x;
    int32_t *x_dcf_pairs, *x_fr_pairs, *x_rest_id;
    uint8_t *ss_map;          /* [nmp*nmp] pairs allowed by a CT/BPSEQ/BPL file, or NULL */
    /* annealing: schedule, current entry (1-based), next switch step, and
     * the kT of the force's limit MC (var_state's kT, never updated) */
    int an_n, ianneal; int64_t *an_step; double *an_T; int64_t an_next; double mc_kT;
    /* timed Rg bias (src/job_md.F90:231-265,614-627) */
    int rgs_n, irgs; int64_t *rgs_step, rgs_next; double *rgs_k, *rgs_0;
    /* box resizing (src/pbc.F90:79-138) */
    int rz_on, rz_has_target; int64_t rz_step, rz_count[3]; double rz_change[3], rz_target[3], rz_orig[3];
    double *x_rg0, *x_dcf_f, *x_fr_k, *x_fr_init, *x_fr_velo, *x_rest_rcut, *x_rest_para;
};

/* ---------------- PBC (src/pbc.F90:140-165, 196-226) ---------------- */
static inline void pbc_vec_d(const orc_sys *s, const double *a, const double *b, double *v) {
    for (int k = 0; k < 3; k++) v[k] = a[k] - b[k];
    if (!s->p.pbc) return;
    for (int k = 0; k < 3; k++) {
        double half = 0.5 * s->p.box[k];
        if (v[k] > half) v[k] = v[k] - s->p.box[k];
        else if (v[k] < -half) v[k] = v[k] + s->p.box[k];
    }
}
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void pbc_wrap(orc_sys *s, int irep) {
    double *x = s->xyz + (size_t)irep * s->nmp * 3;
    for (int i = 0; i < s->nmp; i++)
        for (int k = 0; k < 3; k++) {
            double v = x[3 * i + k], L = s->p.box[k];
            if (v > L) x[3 * i + k] = v - L * ((double)(int)((v - L) / L) + 1);
            else if (v < 0.0) x[3 * i + k] = v + L * ((double)(int)((0.0 - v) / L) + 1);
        }
}

/* ---------------- base-pair tables (src/init_bp.F90:39-172) ---------------- */
static int is_complement(int a, int b) {   /* src/const_idx.F90:339-369 */
    if (a == A_) return b == U_;
    if (a == U_) return b == A_ || b == G_;
    if (a == G_) return b == C_ || b == U_;
    if (a == C_) return b == G_;
    return 0;
}
/* NN stack index for 5'-w x-3' / 3'-z y-5' (src/const_idx.F90:230-330) */
static int seqt2nnt(int w, int x, int z, int y) {
    static const char ch[5] = {'A', 'U', 'G', 'C', 'D'};
    static const char *names[21][2] = {
        {"GC_CG", 0}, {"CC_GG", "GG_CC"}, {"GA_CU", "UC_AG"}, {"CG_GC", 0}, {"AC_UG", "GU_CA"},
        {"CA_GU", "UG_AC"}, {"AG_UC", "CU_GA"}, {"UA_AU", 0}, {"AU_UA", 0}, {"AA_UU", "UU_AA"},
        {"GC_UG", "GU_CG"}, {"CU_GG", "GG_UC"}, {"GG_CU", "UC_GG"}, {"CG_GU", "UG_GC"},
        {"AU_UG", "GU_UA"}, {"GA_UU", "UU_AG"}, {"UG_GU", 0}, {"UA_GU", "UG_AU"},
        {"GG_UU", "UU_GG"}, {"GU_UG", 0}, {"AG_UU", "UU_GA"}};
    char key[6] = {ch[w], ch[x], '_', ch[z], ch[y], 0};
    for (int i = 0; i < 21; i++)
        for (int k = 0; k < 2; k++)
            if (names[i][k] && strcmp(names[i][k], key) == 0) return i + 1;
    return 0;
}
static int bp3_hash_key(int w, int x, int u, int z, int y, int v) {   /* src/init_bp.F90:352-378 */
    int a, b, c, d, e, f;
    if (x < y) { a = w; b = x; c = u; d = z; e = y; f = v; }
    else { f = w; e = x; d = u; c = z; b = y; a = v; }
    return 1 + a + b * 4 + c * 16 + d * 64 + e * 256 + f * 1024;
}

static void nnend_add(orc_sys *s, int h, int cx, int cy, int ox, int oy) {
    /* end correction for centre pair (cx,cy) stacked on outer pair (ox,oy)
     * (src/init_bp.F90:98-167) */
    const sis_params *p = &s->p;
    int au = (cx == A_ && cy == U_) || (cx == U_ && cy == A_);
    int gu = (cx == G_ && cy == U_) || (cx == U_ && cy == G_);
    int oau = (ox == A_ && oy == U_) || (ox == U_ && oy == A_);
    int ogc = (ox == G_ && oy == C_) || (ox == C_ && oy == G_);
    int ogu = (ox == G_ && oy == U_) || (ox == U_ && oy == G_);
    int idx = 0;
    if (au) idx = oau ? 1 : ogc ? 2 : ogu ? 3 : 0;
    else if (gu) idx = oau ? 5 : ogc ? 4 : ogu ? 6 : 0;
    if (idx) {
        s->bp3_dH[h] = s->bp3_dH[h] + p->nnend_dH[idx - 1] - p->bp_dHend0;
        s->bp3_dS[h] = s->bp3_dS[h] + (p->nnend_dS[idx - 1] - p->bp_dSend0) * 1.0e-3;
    }
}

static void init_bp_tables(orc_sys *s) {
    const sis_params *p = &s->p;
    memset(s->bp3_hash, 0, sizeof s->bp3_hash);
    memset(s->bp3_dH, 0, sizeof s->bp3_dH);
    memset(s->bp3_dS, 0, sizeof s->bp3_dS);
    int h = 0;
    /* the table exists for model 5 only (src/init_bp.F90:39) */
    for (int x = 0; x <= 2 && p->bp_model == 5; x++)
        for (int y = x + 1; y <= 3; y++) {
            if (!is_complement(x, y)) continue;
            for (int w = 0; w <= 3; w++)
                for (int z = 0; z <= 3; z++) {
                    int comp_wz = is_complement(w, z);
                    double dH = 0.0, dS = 0.0;
                    if (comp_wz) {
                        dH = 0.5 * (p->nn_dH[seqt2nnt(w, x, z, y) - 1] - p->bp_dH0);
                        dS = 0.5 * (p->nn_dS[seqt2nnt(w, x, z, y) - 1] - p->bp_dS0);
                    }
                    for (int u = 0; u <= 3; u++)
                        for (int v = 0; v <= 3; v++) {
                            int comp_uv = is_complement(u, v);
                            if (!comp_wz && !comp_uv) continue;
                            h++;
                            s->bp3_hash[bp3_hash_key(w, x, u, z, y, v)] = h;
                            if (comp_uv) {
                                s->bp3_dH[h] = dH + 0.5 * (p->nn_dH[seqt2nnt(x, u, y, v) - 1] - p->bp_dH0);
                                s->bp3_dS[h] = (dS + 0.5 * (p->nn_dS[seqt2nnt(x, u, y, v) - 1] - p->bp_dS0)) * 1.0e-3;
                            } else {
                                s->bp3_dH[h] = dH;
                                s->bp3_dS[h] = dS * 1.0e-3;
                            }
                            if (p->nnend) {
                                if (comp_wz && !comp_uv) nnend_add(s, h, x, y, w, z);
                                else if (!comp_wz && comp_uv) nnend_add(s, h, x, y, u, v);
                            }
                        }
                }
        }
    /* BP cutoffs (src/init_bp.F90:321-346) */
    if (fabs(p->bp_cutoff_energy) <= 2.220446049250313e-16) {
        s->bp_cutoff_dist = 18.0;
        for (int t = 1; t <= 6; t++) s->bpp[t].cutoff_ddist = s->bp_cutoff_dist - 13.8;
    } else {
        s->bp_cutoff_dist = 0.0;
        for (int t = 1; t <= 6; t++) {
            s->bpp[t].cutoff_ddist = sqrt(log(fabs(10.0 / p->bp_cutoff_energy)) / s->bpp[t].bond_k);
            double c = s->bpp[t].bond_r + s->bpp[t].cutoff_ddist;
            if (c > s->bp_cutoff_dist) s->bp_cutoff_dist = c;
        }
    }
}

/* Base-pair eligibility of beads imp < jmp (0-based) and its triplet index:
 * the body of the pair loop of src/init_bp.F90:184-313 for models 3/4/5;
 * with a secondary-structure file only the pairs it allows are considered
 * (lines 206-209).  Returns 0 when not eligible. */
void orc_bp_pair(orc_sys *s, int32_t imp, int32_t jmp, int32_t *bptype, int32_t *bp3) {
    *bptype = 0; *bp3 = 0;
    if (imp >= jmp) return;
    if (s->ss_map && !s->ss_map[(size_t)imp * s->nmp + jmp]) return;
    int ich = s->chain_of[imp], jch = s->chain_of[jmp];
    int i = s->lmp[imp], j = s->lmp[jmp];                 /* 1-based local */
    int ni = s->nmp_chain[ich], nj = s->nmp_chain[jch];
    int icirc = s->moltype[ich] == 2, jcirc = s->moltype[jch] == 2;
    int i_pre = i - 1, i_nxt = i + 1, j_pre = j - 1, j_nxt = j + 1;
    if (icirc) { if (i == 1) i_pre = ni; if (i == ni) i_nxt = 1; }
    else { if (i == 1 || i == ni) return; }
    if (jcirc) { if (j == 1) j_pre = nj; if (j == nj) j_nxt = 1; }
    else { if (j == 1 || j == nj) return; }
    const int *si = s->seq + s->chain_start[ich] - 1;   /* si[k], k 1-based */
    const int *sj = s->seq + s->chain_start[jch] - 1;
    int p_ipre_jnxt = is_complement(si[i_pre], sj[j_nxt]);
    int p_inxt_jpre = is_complement(si[i_nxt], sj[j_pre]);
    if (!icirc && i_pre == 1) p_ipre_jnxt = 0;
    if (!jcirc && j_nxt == nj) p_ipre_jnxt = 0;
    if (!icirc && i_nxt == ni) p_inxt_jpre = 0;
    if (!jcirc && j_pre == 1) p_inxt_jpre = 0;
    int L = s->p.bp_min_loop;
    if (ich == jch) {
        if (i + L >= j) return;
        if (i + 1 + L >= j - 1) p_inxt_jpre = 0;
        if (icirc) {
            if (j + L >= i + ni) return;
            if (j + 1 + L >= i - 1 + ni) p_ipre_jnxt = 0;
        }
    }
    int model = s->p.bp_model;
    if ((model == 3 || model == 4 || model == 5) && !p_ipre_jnxt && !p_inxt_jpre) return;
    int a = si[i], b = sj[j], t = 0;
    if (a == G_ && b == C_) t = GC; else if (a == C_ && b == G_) t = CG;
    else if (a == A_ && b == U_) t = AU; else if (a == U_ && b == A_) t = UA;
    else if (a == G_ && b == U_) t = GU; else if (a == U_ && b == G_) t = UG;
    *bptype = t;
    if (model == 5) {
        int key;
        if (!p_ipre_jnxt) key = bp3_hash_key(U_, si[i], si[i_nxt], U_, sj[j], sj[j_pre]);
        else if (!p_inxt_jpre) key = bp3_hash_key(si[i_pre], si[i], U_, sj[j_nxt], sj[j], U_);
        else key = bp3_hash_key(si[i_pre], si[i], si[i_nxt], sj[j_nxt], sj[j], sj[j_pre]);
        *bp3 = s->bp3_hash[key];
    } else {
        *bp3 = t;
    }
}

/* ---------------- set_ele (src/set_ele.F90:45-99) ---------------- */
static void set_ele(orc_sys *s, int irep, double tempk, double ionstr) {
    const double MM_A = 87.740, MM_B = -0.40008, MM_C = 9.398e-4, MM_D = -1.410e-6;
    rep_t *r = &s->rep[irep];
    double Tc = tempk - 273.15;
    r->diele = MM_A + MM_B * Tc + MM_C * Tc * Tc + MM_D * Tc * Tc * Tc;
    double lb = ELE * ELE / (4.0 * PI_ * EPS0 * r->diele * BOLTZ_J * tempk) * 1.0e10;
    double Zp = s->p.length_per_charge / lb;
    r->ele_coef = JOUL2KCAL_MOL * 1.0e10 * (ELE * ELE) / (4.0 * PI_ * EPS0 * r->diele);
    r->ele_coef = r->ele_coef * (Zp * Zp);
    r->lambdaD = 1.0e10 * sqrt((1.0e-3 * EPS0 * r->diele * BOLTZ_J) / (2.0 * N_AVO * (ELE * ELE))) *
                 sqrt(tempk / ionstr);
    if (s->plant_cut_at_set_ele && s->p.ele_cutoff_type == 2) r->ele_cutoff = s->p.ele_cutoff * r->lambdaD;
}

/* Test hook: which = 1 makes set_ele move the cutoff at once (not the
 * reference's behaviour; the planted mistake the ion/temperature ladder
 * fixtures must catch). */
void orc_plant(orc_sys *s, int32_t which, int32_t on) {
    if (which == 1) s->plant_cut_at_set_ele = on;
}

/* set_ele at set-up, followed by init_ele's cutoff (src/init_ele.F90:73-101).
 * Later calls of set_ele (exchange, energy_replica) leave ele_cutoff alone:
 * it changes only in neighbor_list (src/neighbor_list.F90:80-82). */
static void init_ele(orc_sys *s, int irep, double tempk, double ionstr) {
    rep_t *r = &s->rep[irep];
    set_ele(s, irep, tempk, ionstr);
    if (s->p.ele_cutoff_type == 1) r->ele_cutoff = s->p.ele_cutoff;
    else r->ele_cutoff = s->p.ele_cutoff * r->lambdaD;
}

/* ---------------- counter-based random numbers ----------------
 * The engine's performance form draws its normals and exchange uniforms
 * from Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11): ten rounds of the
 * 4x32 S-P network with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85.  Restated here from that paper so the
 * oracle can run the same trajectories; it is not part of the reference. */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        uint64_t a = (uint64_t)0xD2511F53u * c0, b = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0, n1 = (uint32_t)b;
        uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1, n3 = (uint32_t)a;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
/* 53-bit uniform in (0,1) from two words: bits 31..0 of a, then the top
 * 21 bits of b, centred in their interval */
static double ctr_u01(uint32_t a, uint32_t b) {
    uint64_t m = (((uint64_t)a << 21) ^ (uint64_t)(b >> 11)) & ((1ull << 53) - 1);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}
/* three normals for bead `a` from two Philox blocks (counter words c0,c1
 * and tags t0,t1 in the last word): the Box-Muller transform of four
 * uniforms, the pair's sine left over */
static void ctr_normals3(uint32_t c0, uint32_t c1, uint32_t a, uint32_t t0, uint32_t t1,
                         const uint32_t key[2], int init, double n[3]) {
    uint32_t o[4], o2[4];
    uint32_t q[4] = {c0, c1, a, t0}, q2[4] = {init ? c0 - 1u : c0, c1, a, t1};
    orc_philox(q, key, o);
    orc_philox(q2, key, o2);
    double u1 = ctr_u01(o[0], o[1]), u2 = ctr_u01(o[2], o[3]), u3 = ctr_u01(o2[0], o2[1]), u4 = ctr_u01(o2[2], o2[3]);
    double rr = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    n[0] = rr * cos(6.283185307179586 * u2);
    n[1] = rr * sin(6.283185307179586 * u2);
    n[2] = r2 * cos(6.283185307179586 * u4);
}
static void ctr_key(orc_sys *s, int rk, uint32_t key[2]) {
    key[0] = (uint32_t)s->ctr_seed;
    key[1] = (uint32_t)(s->ctr_seed >> 32) ^ (0x9e3779b9u * (uint32_t)(rk + 1));
}
/* normals of replica r, bead a at step istep (the engine's GJF noise) */
static void ctr_step_normals(orc_sys *s, int r, int a, double n[3]) {
    uint32_t key[2];
    ctr_key(s, r, key);
    ctr_normals3((uint32_t)s->istep, (uint32_t)(s->istep >> 32), (uint32_t)a, 0x6e6f6973u, 0x6e6f6974u, key, 0, n);
}
/* initial velocities (which 0) and noise terms (which 1) */
static void ctr_init_normals(orc_sys *s, int r, int a, int which, double n[3]) {
    uint32_t key[2];
    ctr_key(s, r, key);
    ctr_normals3(0xffffffffu, (uint32_t)which, (uint32_t)a, 0x696e6974u, 0x696e6974u, key, 1, n);
}
/* the k-th exchange uniform of step istep (0-based pair counter) */
static double ctr_exchange_uniform(orc_sys *s, int k) {
    uint32_t o[4];
    uint32_t q[4] = {(uint32_t)s->istep, (uint32_t)(s->istep >> 32), (uint32_t)k, 0x72657863u};
    uint32_t key[2] = {(uint32_t)s->ctr_seed, (uint32_t)(s->ctr_seed >> 32) ^ 0x1234567u};
    orc_philox(q, key, o);
    return ctr_u01(o[0], o[1]);
}
void orc_set_counter_rng(orc_sys *s, uint64_t seed, int32_t on) { s->ctr_on = on; s->ctr_seed = seed; }

/* ---------------- MT19937 words and conversions ---------------- */
/* init_genrand (mt_stream.F90:419-455) and genrand_int32 with tempering
 * (mt_stream.F90:510-549). */
void orc_mt_words(uint32_t seed, int64_t n, uint32_t *out) {
    uint32_t mt[624];
    mt[0] = seed;
    for (int i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    int idx = 624;
    for (int64_t k = 0; k < n; k++) {
        if (idx >= 624) {
            for (int i = 0; i < 624; i++) {
                uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        out[k] = y;
    }
}

static tape_t *tape_of(orc_sys *s, int irep) { return irep < 0 ? &s->tape_rep : &s->rep[irep].tape; }
static uint32_t next_word(tape_t *t) {
    if (t->pos >= t->n) { abort(); }
    return t->w[t->pos++];
}
/* genrand_double1: [0,1], 53-bit (mt_stream.F90:708-729) */
static double gen_d1(tape_t *t) {
    uint32_t ia = next_word(t) >> 5, ib = next_word(t) >> 6;
    double a = (double)ia, b = (double)ib;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740991.0);
}
/* genrand_double3: (0,1), 52-bit (mt_stream.F90:754-775) */
static double gen_d3(tape_t *t) {
    uint32_t ia = next_word(t) >> 6, ib = next_word(t) >> 6;
    double a = (double)ia, b = (double)ib;
    return (a * 67108864.0 + b + 0.5) * (1.0 / 4503599627370496.0);
}
/* rnd_boxmuller (src/rnd_boxmuller.F90:1-49): polar method, one cached
 * value shared by all replicas. */
static double boxmuller(orc_sys *s, int irep) {
    if (!s->bm_stored) {
        s->bm_stored = 1;
        tape_t *t = tape_of(s, irep);
        double vx, vy, r2;
        for (;;) {
            vx = 2.0 * gen_d1(t) - 1.0;
            vy = 2.0 * gen_d1(t) - 1.0;
            r2 = vx * vx + vy * vy;
            if (r2 < 1.0 && r2 > 0.0) break;
        }
        double rf = sqrt(-2.0 * log(r2) / r2);
        s->bm_value = vx * rf;
        return vy * rf;
    }
    s->bm_stored = 0;
    return s->bm_value;
}

/* One (imp, jmp) pair of neighbor_list's loops, chain ich position i,
 * chain jch position j (1-based), in the reference's loop body
 * (src/neighbor_list.F90:119-223). */
typedef struct { double wca2, ele2, bp2; } nl_cuts;
static void nl_pair(orc_sys *s, rep_t *r, const double *x, const nl_cuts *c, int ich, int jch, int i, int j) {
    int imp = s->chain_start[ich] + i - 1, jmp = s->chain_start[jch] + j - 1;
    int iseq = s->seq[imp], jseq = s->seq[jmp];
    double v[3];
    pbc_vec_d(s, x + 3 * imp, x + 3 * jmp, v);
    double d2 = dot3(v, v);
    if (iseq != D_ && jseq != D_ && d2 <= c->wca2) {
        int add = 1;
        if (ich == jch) {
            if (i + 2 >= j) add = 0;
            else if (s->moltype[ich] == 2 && j + 2 >= i + s->nmp_chain[ich]) add = 0;
        }
        if (add) { ipush(&r->wca, imp + 1); ipush(&r->wca, jmp + 1); }
    }
    if (s->p.ele_on && d2 <= c->ele2 && s->has_charge[imp] && s->has_charge[jmp]) {
        int add = 1;
        if (ich == jch) {
            int n = s->nmp_chain[ich];
            if (j == i + 1 && s->p.ele_excl_bond) add = 0;
            if (j == i + 2 && s->p.ele_excl_angl) add = 0;
            if (s->moltype[ich] == 2) {
                if (i == 1) {
                    if (j == n && s->p.ele_excl_bond) add = 0;
                    if (j == n - 1 && s->p.ele_excl_angl) add = 0;
                } else if (i == 2) {
                    if (j == n && s->p.ele_excl_angl) add = 0;
                }
            }
        }
        if (add) { ipush(&r->ele, imp + 1); ipush(&r->ele, jmp + 1); }
    }
    if (d2 > c->bp2) return;
    int bt, b3;
    orc_bp_pair(s, imp, jmp, &bt, &b3);
    if (b3 > 0) {
        int ni = s->nmp_chain[ich], nj = s->nmp_chain[jch];
        ipush(&r->bp, imp + 1);
        ipush(&r->bp, jmp + 1);
        ipush(&r->bp, (i == 1) ? imp + 1 + ni - 1 : imp + 1 - 1);
        ipush(&r->bp, (j == 1) ? jmp + 1 + nj - 1 : jmp + 1 - 1);
        ipush(&r->bp, (i == ni) ? imp + 1 + 1 - ni : imp + 1 + 1);
        /* the reference compares j with nmp_chain(ichain) here
         * (src/neighbor_list.F90:212) */
        ipush(&r->bp, (j == ni) ? jmp + 1 + 1 - nj : jmp + 1 + 1);
        ipush(&r->bp, bt);
        /* the triplet table is zero for models other than 5
         * (src/init_bp.F90:39, src/neighbor_list.F90:220-221) */
        dpush(&r->bpcoef, s->p.bp_model == 5 ? s->bp3_dH[b3] : 0.0);
        dpush(&r->bpcoef, s->p.bp_model == 5 ? s->bp3_dS[b3] : 0.0);
    }
}

typedef struct { int ich, jch, a, b; } nl_cand;
/* the loops' order: chain of the first bead, chain of the second, then
 * the positions (bead ids grow along a chain) */
static int cmp_cand(const void *p, const void *q) {
    const nl_cand *x = p, *y = q;
    if (x->ich != y->ich) return x->ich < y->ich ? -1 : 1;
    if (x->jch != y->jch) return x->jch < y->jch ? -1 : 1;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    return (x->b > y->b) - (x->b < y->b);
}

/* Cell-list enumeration of the same pairs for large periodic systems: the
 * reference's O(N^2) loops visit every pair in (ichain, jchain >= ichain,
 * i, j) order and test distances; pairs farther apart than the largest
 * list cutoff add nothing, so the candidates found through cells of at
 * least that width (27 neighbours, minimum image) are sorted into the
 * loops' order and given the same loop body.  The result equals the
 * brute-force lists (pinned by test_oracle.py on periodic systems). */
static int nl_cells(orc_sys *s, rep_t *r, const double *x, const nl_cuts *c) {
    double cut2 = c->wca2;
    if (c->ele2 > cut2) cut2 = c->ele2;
    if (c->bp2 > cut2) cut2 = c->bp2;
    const double cut = sqrt(cut2);
    int nc[3];
    for (int k = 0; k < 3; k++) {
        nc[k] = (int)floor(s->p.box[k] / cut);
        if (nc[k] < 3) return 0;
        if (nc[k] > 256) nc[k] = 256;
    }
    const int n = s->nmp;
    const int ncell = nc[0] * nc[1] * nc[2];
    int *cell = malloc(sizeof(int) * n), *cnt = calloc(ncell + 1, sizeof(int)), *beads = malloc(sizeof(int) * n);
    for (int a = 0; a < n; a++) {
        int q[3];
        for (int k = 0; k < 3; k++) {
            q[k] = (int)(x[3 * a + k] * nc[k] / s->p.box[k]);
            if (q[k] < 0) q[k] = 0;
            if (q[k] >= nc[k]) q[k] = nc[k] - 1;
        }
        cell[a] = (q[2] * nc[1] + q[1]) * nc[0] + q[0];
        cnt[cell[a] + 1]++;
    }
    for (int k = 0; k < ncell; k++) cnt[k + 1] += cnt[k];
    int *fill = malloc(sizeof(int) * ncell);
    memcpy(fill, cnt, sizeof(int) * ncell);
    for (int a = 0; a < n; a++) beads[fill[cell[a]]++] = a;
    free(fill);
    /* candidate pairs, each once: the loops take a pair (a, b) with a in
     * the earlier chain, or earlier in the same chain, i.e. a < b */
    int64_t cap = 1 << 20, m = 0;
    nl_cand *key = malloc(sizeof(nl_cand) * cap);
    for (int a = 0; a < n; a++) {
        int cx = cell[a] % nc[0], cy = (cell[a] / nc[0]) % nc[1], cz = cell[a] / (nc[0] * nc[1]);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int qx = (cx + dx + nc[0]) % nc[0], qy = (cy + dy + nc[1]) % nc[1], qz = (cz + dz + nc[2]) % nc[2];
                    int q = (qz * nc[1] + qy) * nc[0] + qx;
                    for (int t = cnt[q]; t < cnt[q + 1]; t++) {
                        int b = beads[t];
                        if (b <= a) continue;
                        double v[3];
                        pbc_vec_d(s, x + 3 * a, x + 3 * b, v);
                        if (dot3(v, v) > cut2) continue;
                        if (m == cap) { cap *= 2; key = realloc(key, sizeof(nl_cand) * cap); }
                        key[m].ich = s->chain_of[a]; key[m].jch = s->chain_of[b];
                        key[m].a = a; key[m].b = b;
                        m++;
                    }
                }
    }
    /* with at least 3 cells per axis the 27 neighbour cells are distinct,
     * so every pair is found once */
    qsort(key, m, sizeof(nl_cand), cmp_cand);
    for (int64_t k = 0; k < m; k++)
        nl_pair(s, r, x, c, key[k].ich, key[k].jch, s->lmp[key[k].a], s->lmp[key[k].b]);
    free(key); free(cell); free(cnt); free(beads);
    return 1;
}

/* ---------------- neighbour lists (src/neighbor_list.F90:84-233) ---------------- */
void orc_neighbor_list(orc_sys *s, int32_t irep1) {
    int irep = irep1 - 1;
    rep_t *r = &s->rep[irep];
    if (s->p.pbc) pbc_wrap(s, irep);
    r->wca.n = r->ele.n = r->bp.n = r->bpcoef.n = 0;
    const double margin = s->p.nl_margin;
    nl_cuts c;
    c.ele2 = 0.0;
    if (s->p.ele_on && s->p.ele_cutoff_type == 2) r->ele_cutoff = s->p.ele_cutoff * r->lambdaD;
    if (s->p.ele_on) c.ele2 = (r->ele_cutoff + margin) * (r->ele_cutoff + margin);
    c.wca2 = (s->p.wca_sigma + margin) * (s->p.wca_sigma + margin);
    c.bp2 = (s->bp_cutoff_dist + margin) * (s->bp_cutoff_dist + margin);
    const double *x = s->xyz + (size_t)irep * s->nmp * 3;
    if (s->p.pbc && s->nmp >= 4096 && !s->nl_brute && nl_cells(s, r, x, &c)) goto done;
    for (int ich = 0; ich < s->nchains; ich++)
        for (int jch = ich; jch < s->nchains; jch++)
            for (int i = 1; i <= s->nmp_chain[ich]; i++) {
                int j_start = (ich == jch) ? i + 1 : 1;
                for (int j = j_start; j <= s->nmp_chain[jch]; j++) nl_pair(s, r, x, &c, ich, jch, i, j);
            }
done:;
    int64_t nbp = r->bp.n / 7;
    r->bp_status = realloc(r->bp_status, nbp + 1);
    r->ene_bp = realloc(r->ene_bp, (nbp + 1) * sizeof(double));
    memset(r->bp_status, 0, nbp + 1);
    memset(r->ene_bp, 0, (nbp + 1) * sizeof(double));
    r->bp_status_mc = realloc(r->bp_status_mc, nbp + 1);
    memset(r->bp_status_mc, 0, nbp + 1);
    r->have_status_mc = 0;
}

void orc_set_nl_brute(orc_sys *s, int32_t on) { s->nl_brute = on; }

/* The cutoff-free lists of the DCD and CHECK_FORCE jobs, for replica irep:
 * list_exv (src/list_exv.F90:19-64: j >= i + 3 within a chain, no circular
 * wrap rule), list_ele (src/list_ele.F90:18-65: bond / angle exclusions
 * without the circular wrap) in (chain i, chain j, i, j) order, and
 * list_bp (src/list_bp.F90:41-93) in global (imp, jmp) order with
 * jmp >= imp + min_loop + 1 over the whole system (also across chains),
 * the neighbours' wrap taken only for circular chains, each with the
 * bp3_map type and triplet coefficients. */
void orc_analysis_lists(orc_sys *s, int32_t irep1) {
    rep_t *r = &s->rep[irep1 - 1];
    r->wca.n = r->ele.n = r->bp.n = r->bpcoef.n = 0;
    /* init_ele's cutoff (src/init_ele.F90:72-104) */
    if (s->p.ele_on && s->p.ele_cutoff_type == 2) r->ele_cutoff = s->p.ele_cutoff * r->lambdaD;
    for (int ich = 0; ich < s->nchains; ich++)
        for (int jch = ich; jch < s->nchains; jch++)
            for (int i = 1; i <= s->nmp_chain[ich]; i++) {
                int imp = s->chain_start[ich] + i - 1;
                if (s->seq[imp] == D_) continue;
                for (int j = (ich == jch) ? i + 3 : 1; j <= s->nmp_chain[jch]; j++) {
                    int jmp = s->chain_start[jch] + j - 1;
                    if (s->seq[jmp] == D_) continue;
                    ipush(&r->wca, imp + 1); ipush(&r->wca, jmp + 1);
                }
            }
    if (s->p.ele_on)
        for (int ich = 0; ich < s->nchains; ich++)
            for (int jch = ich; jch < s->nchains; jch++)
                for (int i = 1; i <= s->nmp_chain[ich]; i++) {
                    int imp = s->chain_start[ich] + i - 1;
                    if (!s->has_charge[imp]) continue;
                    for (int j = (ich == jch) ? i + 1 : 1; j <= s->nmp_chain[jch]; j++) {
                        /* the reference tests j against i for every chain pair
                         * (src/list_ele.F90:45-46) */
                        if (j == i + 1 && s->p.ele_excl_bond) continue;
                        if (j == i + 2 && s->p.ele_excl_angl) continue;
                        int jmp = s->chain_start[jch] + j - 1;
                        if (!s->has_charge[jmp]) continue;
                        ipush(&r->ele, imp + 1); ipush(&r->ele, jmp + 1);
                    }
                }
    for (int imp = 0; imp < s->nmp; imp++)
        for (int jmp = imp + s->p.bp_min_loop + 1; jmp < s->nmp; jmp++) {
            int bt, b3;
            orc_bp_pair(s, imp, jmp, &bt, &b3);
            if (b3 == 0) continue;
            int ich = s->chain_of[imp], jch = s->chain_of[jmp];
            int ni = s->nmp_chain[ich], nj = s->nmp_chain[jch];
            int ic = s->moltype[ich] == 2, jc = s->moltype[jch] == 2;
            int li = s->lmp[imp], lj = s->lmp[jmp];
            ipush(&r->bp, imp + 1);
            ipush(&r->bp, jmp + 1);
            ipush(&r->bp, (ic && li == 1) ? imp + 1 + ni - 1 : imp + 1 - 1);
            ipush(&r->bp, (jc && lj == 1) ? jmp + 1 + nj - 1 : jmp + 1 - 1);
            ipush(&r->bp, (ic && li == ni) ? imp + 1 - ni + 1 : imp + 1 + 1);
            ipush(&r->bp, (jc && lj == nj) ? jmp + 1 - nj + 1 : jmp + 1 + 1);
            ipush(&r->bp, bt);
            dpush(&r->bpcoef, s->p.bp_model == 5 ? s->bp3_dH[b3] : 0.0);
            dpush(&r->bpcoef, s->p.bp_model == 5 ? s->bp3_dS[b3] : 0.0);
        }
    int64_t nbp = r->bp.n / 7;
    r->bp_status = realloc(r->bp_status, nbp + 1);
    r->ene_bp = realloc(r->ene_bp, (nbp + 1) * sizeof(double));
    memset(r->bp_status, 0, nbp + 1);
    memset(r->ene_bp, 0, (nbp + 1) * sizeof(double));
    r->bp_status_mc = realloc(r->bp_status_mc, nbp + 1);
    memset(r->bp_status_mc, 0, nbp + 1);
    r->have_status_mc = 0;
}

int64_t orc_count(orc_sys *s, int32_t irep, int32_t kind) {
    rep_t *r = &s->rep[irep - 1];
    return kind == 0 ? r->wca.n / 2 : kind == 1 ? r->ele.n / 2 : r->bp.n / 7;
}
void orc_lists(orc_sys *s, int32_t irep, int32_t kind, int32_t *out, double *bpcoef) {
    rep_t *r = &s->rep[irep - 1];
    ivec *v = kind == 0 ? &r->wca : kind == 1 ? &r->ele : &r->bp;
    memcpy(out, v->a, v->n * sizeof(int32_t));
    if (kind == 2 && bpcoef) memcpy(bpcoef, r->bpcoef.a, r->bpcoef.n * sizeof(double));
}

/* ---------------- forces ---------------- */
#define F3(f, i) (f + 3 * (size_t)(i))
static void force_bond(orc_sys *s, const double *x, double *f) {    /* src/force_bond.F90:33-57 */
    for (int c = 0; c < s->nchains; c++) {
        int n = s->nmp_chain[c], st = s->chain_start[c];
        int nb = n - 1 + (s->moltype[c] == 2);
        for (int b = 0; b < nb; b++) {
            int i1 = (b < n - 1) ? st + b : st + n - 1, i2 = (b < n - 1) ? st + b + 1 : st;
            double v[3], fv[3];
            pbc_vec_d(s, F3(x, i1), F3(x, i2), v);
            double d = sqrt(dot3(v, v));
            double delta = d - s->p.bond_r0;
            for (int k = 0; k < 3; k++) fv[k] = s->p.bond_k * delta / d * v[k];
            for (int k = 0; k < 3; k++) { F3(f, i1)[k] -= fv[k]; F3(f, i2)[k] += fv[k]; }
        }
    }
}
/* angle / dihedral index lists in list_local order (src/list_local.F90:24-104) */
static int angle_list(orc_sys *s, int (*out)[3]) {
    int n = 0;
    for (int c = 0; c < s->nchains; c++) {
        int m = s->nmp_chain[c], st = s->chain_start[c];
        for (int i = 1; i <= m - 1; i++)
            if (i != m - 1) { out[n][0] = st + i - 1; out[n][1] = st + i; out[n][2] = st + i + 1; n++; }
        if (s->moltype[c] == 2) {
            int ini = st, end = st + m - 1;
            out[n][0] = end - 1; out[n][1] = end; out[n][2] = ini; n++;
            out[n][0] = end; out[n][1] = ini; out[n][2] = ini + 1; n++;
        }
    }
    return n;
}
static int dih_list(orc_sys *s, int (*out)[4]) {
    int n = 0;
    for (int c = 0; c < s->nchains; c++) {
        int m = s->nmp_chain[c], st = s->chain_start[c];
        for (int i = 1; i <= m - 1; i++)
            if (i < m - 2) { for (int k = 0; k < 4; k++) out[n][k] = st + i - 1 + k; n++; }
        if (s->moltype[c] == 2) {
            int ini = st, end = st + m - 1;
            int q[3][4] = {{end - 2, end - 1, end, ini}, {end - 1, end, ini, ini + 1}, {end, ini, ini + 1, ini + 2}};
            for (int a = 0; a < 3; a++) { for (int k = 0; k < 4; k++) out[n][k] = q[a][k]; n++; }
        }
    }
    return n;
}
static void force_angl(orc_sys *s, const double *x, double *f) {
    int (*al)[3] = malloc(sizeof(int[3]) * (s->nmp + 2 * s->nchains));
    int na = angle_list(s, al);
    const double k = s->p.angl_k, t0 = s->p.angl_t0;
    for (int a = 0; a < na; a++) {
        int i1 = al[a][0], i2 = al[a][1], i3 = al[a][2];
        if (s->p.angl_reb) {   /* src/force_angl_ReB.F90:41-74 */
            double co0 = cos(t0), v12[3], v32[3], f1[3], f3[3];
            pbc_vec_d(s, F3(x, i1), F3(x, i2), v12);
            pbc_vec_d(s, F3(x, i3), F3(x, i2), v32);
            double d12 = dot3(v12, v12), d32 = dot3(v32, v32), a1232 = sqrt(d12 * d32);
            double cosine = dot3(v12, v32) / a1232;
            double den = (1.0 - cosine * cosine);
            double ff = -k * (cosine - co0) * (1.0 - cosine * co0) / (den * den);
            for (int c = 0; c < 3; c++) {
                f1[c] = ff * (v32[c] / a1232 - cosine * v12[c] / d12);
                f3[c] = ff * (v12[c] / a1232 - cosine * v32[c] / d32);
            }
            for (int c = 0; c < 3; c++) {
                F3(f, i1)[c] += f1[c]; F3(f, i2)[c] -= (f1[c] + f3[c]); F3(f, i3)[c] += f3[c];
            }
        } else {               /* src/force_angl.F90:35-83 */
            double v21[3], v32[3], f21[3], f32[3];
            pbc_vec_d(s, F3(x, i2), F3(x, i1), v21);
            pbc_vec_d(s, F3(x, i3), F3(x, i2), v32);
            double d21 = dot3(v21, v21), d32 = dot3(v32, v32), d2132 = dot3(v32, v21);
            double cosine = -d2132 / sqrt(d21 * d32);
            if (fabs(cosine) > 1.0) cosine = copysign(1.0, cosine);
            double delta = acos(cosine) - t0;
            double t = d21 * d32 - d2132 * d2132;
            if (t <= 1.0) t = 1.0;
            double ff = k * delta / sqrt(t);
            for (int c = 0; c < 3; c++) {
                f21[c] = ff * (v21[c] * (d2132 / d21) - v32[c]);
                f32[c] = ff * (v32[c] * (d2132 / d32) - v21[c]);
            }
            for (int c = 0; c < 3; c++) {
                F3(f, i1)[c] -= f21[c]; F3(f, i2)[c] += f21[c] - f32[c]; F3(f, i3)[c] += f32[c];
            }
        }
    }
    free(al);
}
static void cross(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static void force_dih(orc_sys *s, const double *x, double *f) {
    /* src/force_dih_exp.F90:36-95 and src/force_dih_cos.F90:21-60 */
    if (s->p.dih_type == 0) return;
    int (*dl)[4] = malloc(sizeof(int[4]) * (s->nmp + 3 * s->nchains));
    int nd = dih_list(s, dl);
    double coef1 = s->p.dih_k * s->p.dih_w, coef2 = -0.5 * s->p.dih_w;
    for (int a = 0; a < nd; a++) {
        int i1 = dl[a][0], i2 = dl[a][1], i3 = dl[a][2], i4 = dl[a][3];
        double vij[3], vkj[3], vkl[3], m[3], n[3], fi[3], fj[3], fk[3], fl[3];
        pbc_vec_d(s, F3(x, i1), F3(x, i2), vij);
        pbc_vec_d(s, F3(x, i3), F3(x, i2), vkj);
        pbc_vec_d(s, F3(x, i3), F3(x, i4), vkl);
        cross(vij, vkj, m);
        cross(vkj, vkl, n);
        double akj2 = dot3(vkj, vkj), akj = sqrt(akj2);
        double dih = atan2(akj * dot3(vij, n), dot3(m, n));
        double pre;
        if (s->p.dih_type == 2) {
            double dd = dih - s->p.dih_p0;
            pre = coef1 * dd * exp(coef2 * (dd * dd)) * akj;
        } else {
            pre = -s->p.dih_k * sin(dih + s->p.dih_p0) * akj;
        }
        double mm = dot3(m, m), nn = dot3(n, n);
        for (int c = 0; c < 3; c++) { fi[c] = -pre / mm * m[c]; fl[c] = pre / nn * n[c]; }
        double dvij = dot3(vij, vkj) / akj2, dvkl = dot3(vkl, vkj) / akj2;
        for (int c = 0; c < 3; c++) {
            fj[c] = (-1.0 + dvij) * fi[c] + (-dvkl) * fl[c];
            fk[c] = (-dvij) * fi[c] + (-1.0 + dvkl) * fl[c];
        }
        for (int c = 0; c < 3; c++) {
            F3(f, i1)[c] += fi[c]; F3(f, i2)[c] += fj[c]; F3(f, i3)[c] += fk[c]; F3(f, i4)[c] += fl[c];
        }
    }
    free(dl);
}
static void force_wca(orc_sys *s, int irep, const double *x, double *f) {   /* src/force_wca.F90:34-67 */
    rep_t *r = &s->rep[irep];
    double sig2 = s->p.wca_sigma * s->p.wca_sigma, coef = 12.0 * s->p.wca_eps / sig2;
    for (int64_t k = 0; k < r->wca.n / 2; k++) {
        int i1 = r->wca.a[2 * k] - 1, i2 = r->wca.a[2 * k + 1] - 1;
        double v[3];
        pbc_vec_d(s, F3(x, i1), F3(x, i2), v);
        double d2 = dot3(v, v);
        if (d2 >= sig2) continue;
        d2 = sig2 / d2;
        double d4 = d2 * d2 * d2 * d2, d7 = d4 * d2 * d2 * d2;
        for (int c = 0; c < 3; c++) { double fv = coef * (d7 - d4) * v[c]; F3(f, i1)[c] += fv; F3(f, i2)[c] -= fv; }
    }
}
static void force_ele(orc_sys *s, int irep, const double *x, double *f) {  /* src/force_ele_DH.F90:39-82 */
    rep_t *r = &s->rep[irep];
    double cutoff2 = r->ele_cutoff * r->ele_cutoff, rcdist = 1.0 / r->lambdaD;
    for (int64_t k = 0; k < r->ele.n / 2; k++) {
        int i1 = r->ele.a[2 * k] - 1, i2 = r->ele.a[2 * k + 1] - 1;
        double v[3];
        pbc_vec_d(s, F3(x, i2), F3(x, i1), v);
        double dist2 = dot3(v, v);
        if (dist2 > cutoff2) continue;
        double dist1 = sqrt(dist2), rdist1 = 1.0 / dist1;
        double dv = r->ele_coef * rdist1 * rdist1 * (rdist1 + rcdist) * exp(-dist1 * rcdist);
        for (int c = 0; c < 3; c++) { double fv = dv * v[c]; F3(f, i1)[c] -= fv; F3(f, i2)[c] += fv; }
    }
}

/* 6-bead base-pair geometry: energy factor u and unscaled forces fb[6][3]
 * (src/force_bp_limit_triplet.F90:100-240).  Returns 0 when the pair is
 * outside cutoff_ddist. */
static int bp_geom(orc_sys *s, const double *x, const int32_t *mp, double *u_out, double fb[6][3]) {
    const sis_bp_para *bpp = &s->bpp[mp[6]];
    int i1 = mp[0] - 1, i2 = mp[1] - 1, i3 = mp[2] - 1, i4 = mp[3] - 1, i5 = mp[4] - 1, i6 = mp[5] - 1;
    double v12[3], v13[3], v15[3], v42[3], v62[3], fi[3], fk[3], m[3], n[3];
    pbc_vec_d(s, F3(x, i1), F3(x, i2), v12);
    double d1212 = dot3(v12, v12), a12 = sqrt(d1212), d = a12 - bpp->bond_r;
    if (fabs(d) > bpp->cutoff_ddist) return 0;
    double u = bpp->bond_k * (d * d);
    for (int c = 0; c < 3; c++) { fi[c] = (2.0 * bpp->bond_k * d / a12) * v12[c]; fb[0][c] = fi[c]; fb[1][c] = -fi[c]; }
    pbc_vec_d(s, F3(x, i1), F3(x, i3), v13);
    pbc_vec_d(s, F3(x, i1), F3(x, i5), v15);
    pbc_vec_d(s, F3(x, i4), F3(x, i2), v42);
    pbc_vec_d(s, F3(x, i6), F3(x, i2), v62);
    double d1313 = dot3(v13, v13), d4242 = dot3(v42, v42), d1213 = dot3(v12, v13);
    double d1242 = dot3(v12, v42), d1215 = dot3(v12, v15), d1515 = dot3(v15, v15);
    double d1262 = dot3(v12, v62), d6262 = dot3(v62, v62);
    double d1213o = d1213 / d1212, d1242o = d1242 / d1212, d1215o = d1215 / d1212, d1262o = d1262 / d1212;
    double cosine, pre;
    /* angle 3-1=2 */
    cosine = d1213 / (sqrt(d1313) * a12);
    d = acos(cosine) - bpp->angl_theta2;
    pre = 2.0 * bpp->angl_k2 * d / sqrt(d1313 * d1212 - d1213 * d1213);
    u = u + bpp->angl_k2 * (d * d);
    for (int c = 0; c < 3; c++) {
        fi[c] = pre * (v12[c] - (d1213 / d1313 * v13[c]));
        fk[c] = pre * (v13[c] - (d1213o * v12[c]));
        fb[0][c] = fb[0][c] - fi[c] - fk[c]; fb[1][c] = fb[1][c] + fk[c]; fb[2][c] = fi[c];
    }
    /* angle 1=2-4 */
    cosine = d1242 / (a12 * sqrt(d4242));
    d = acos(cosine) - bpp->angl_theta1;
    u = u + bpp->angl_k1 * (d * d);
    pre = 2.0 * bpp->angl_k1 * d / sqrt(d1212 * d4242 - d1242 * d1242);
    for (int c = 0; c < 3; c++) {
        fi[c] = -pre * (v42[c] - (d1242o * v12[c]));
        fk[c] = -pre * (v12[c] - (d1242 / d4242 * v42[c]));
        fb[0][c] = fb[0][c] + fi[c]; fb[1][c] = fb[1][c] - fi[c] - fk[c]; fb[3][c] = fk[c];
    }
    /* angle 5-1=2 */
    cosine = d1215 / (sqrt(d1515) * a12);
    d = acos(cosine) - bpp->angl_theta4;
    u = u + bpp->angl_k4 * (d * d);
    pre = 2.0 * bpp->angl_k4 * d / sqrt(d1515 * d1212 - d1215 * d1215);
    for (int c = 0; c < 3; c++) {
        fi[c] = pre * (v12[c] - (d1215 / d1515 * v15[c]));
        fk[c] = pre * (v15[c] - (d1215o * v12[c]));
        fb[0][c] = fb[0][c] - fi[c] - fk[c]; fb[1][c] = fb[1][c] + fk[c]; fb[4][c] = fi[c];
    }
    /* angle 1=2-6 */
    cosine = d1262 / (a12 * sqrt(d6262));
    d = acos(cosine) - bpp->angl_theta3;
    u = u + bpp->angl_k3 * (d * d);
    pre = 2.0 * bpp->angl_k3 * d / sqrt(d1212 * d6262 - d1262 * d1262);
    for (int c = 0; c < 3; c++) {
        fi[c] = -pre * (v62[c] - (d1262o * v12[c]));
        fk[c] = -pre * (v12[c] - (d1262 / d6262 * v62[c]));
        fb[0][c] = fb[0][c] + fi[c]; fb[1][c] = fb[1][c] - fi[c] - fk[c]; fb[5][c] = fk[c];
    }
    /* dihedral 4-2=1=3 */
    cross(v42, v12, m);
    cross(v12, v13, n);
    double dih = atan2(dot3(v42, n) * a12, dot3(m, n));
    d = dih + bpp->dihd_phi1;
    u = u + bpp->dihd_k1 * (1.0 + cos(d));
    pre = -bpp->dihd_k1 * sin(d) * a12;
    double mm = dot3(m, m), nn = dot3(n, n);
    double fl[3];
    for (int c = 0; c < 3; c++) {
        fi[c] = +pre / mm * m[c];
        fl[c] = -pre / nn * n[c];
        fb[3][c] = fb[3][c] + fi[c];
        fb[1][c] = fb[1][c] + (-1.0 + d1242o) * fi[c] - (d1213o) * fl[c];
        fb[0][c] = fb[0][c] + (-1.0 + d1213o) * fl[c] - (d1242o) * fi[c];
        fb[2][c] = fb[2][c] + fl[c];
    }
    /* dihedral 6-2=1=5 */
    cross(v62, v12, m);
    cross(v12, v15, n);
    dih = atan2(dot3(v62, n) * a12, dot3(m, n));
    d = dih + bpp->dihd_phi2;
    u = u + bpp->dihd_k2 * (1.0 + cos(d));
    pre = -bpp->dihd_k2 * sin(d) * a12;
    mm = dot3(m, m); nn = dot3(n, n);
    for (int c = 0; c < 3; c++) {
        fi[c] = +pre / mm * m[c];
        fl[c] = -pre / nn * n[c];
        fb[5][c] = fb[5][c] + fi[c];
        fb[1][c] = fb[1][c] + (-1.0 + d1262o) * fi[c] - (d1215o) * fl[c];
        fb[0][c] = fb[0][c] + (-1.0 + d1215o) * fl[c] - (d1262o) * fi[c];
        fb[4][c] = fb[4][c] + fl[c];
    }
    *u_out = u;
    return 1;
}

/* The limit Monte Carlo that removes excess base pairs per nucleotide,
 * shared by force_bp_limit_triplet (src/force_bp_limit_triplet.F90:259-330)
 * and energy_bp_limit_triplet (src/energy_bp_limit_triplet.F90:108-190). */
static int64_t mc_iterations = 0;
int64_t orc_mc_iterations(void) { return mc_iterations; }
static void bp_limit_mc(orc_sys *s, int irep, int *nt_excess, double beta) {
    rep_t *r = &s->rep[irep];
    int64_t nbp = r->bp.n / 7;
    int *ntlist = malloc(sizeof(int) * (s->nmp + 1));
    int *bp_seq = malloc(sizeof(int) * (nbp + 1));
    tape_t *t = &r->tape;
    for (;;) {
        int nnt = 0;
        for (int imp = 0; imp < s->nmp; imp++) if (nt_excess[imp] > 0) ntlist[++nnt] = imp + 1;
        if (nnt == 0) break;
        mc_iterations++;
        double rnd = gen_d3(t);
        int nt_delete = ntlist[(int)ceil(rnd * nnt)];
        int nseq = 0;
        for (int64_t b = 0; b < nbp; b++)
            if (r->bp_status[b]) {
                int i1 = r->bp.a[7 * b], i2 = r->bp.a[7 * b + 1];
                if (i1 == nt_delete || i2 == nt_delete) bp_seq[++nseq] = (int)b;
            }
        for (int i = 1; i <= nseq; i++) {
            rnd = gen_d3(t);
            int isw = (int)ceil(rnd * nseq);
            int sv = bp_seq[i]; bp_seq[i] = bp_seq[isw]; bp_seq[isw] = sv;
        }
        int del = bp_seq[1];
        for (int i = 2; i <= nseq; i++) {
            int jb = bp_seq[i];
            double ratio = exp((r->ene_bp[jb] - r->ene_bp[del]) * beta);
            rnd = gen_d1(t);
            if (rnd < ratio) del = jb;
        }
        r->bp_status[del] = 0;
        nt_excess[r->bp.a[7 * del] - 1] -= 1;
        nt_excess[r->bp.a[7 * del + 1] - 1] -= 1;
    }
    free(ntlist);
    free(bp_seq);
}

/* prefactor of a candidate's energy: dG = dH - T dS of its triplet for
 * model 5 (src/force_bp_triplet.F90:90-97), 1 for models 1/3, whose
 * energy is U0 exp(-u) (src/force_bp.F90:203) */
static double bp_prefactor(const orc_sys *s, const rep_t *r, int64_t b, double tK) {
    if (s->p.bp_model != 5) return 1.0;
    return r->bpcoef.a[2 * b] - tK * r->bpcoef.a[2 * b + 1];
}

static void force_bp(orc_sys *s, int irep, int bp_mc, const double *x, double *f) {
    /* models 4/5: src/force_bp_limit_triplet.F90 (bp_mc) and
     * src/force_bp_triplet.F90 (no MC, skips candidates not in bp_status_MC);
     * models 1/3 without MC: src/force_bp.F90 */
    rep_t *r = &s->rep[irep];
    int64_t nbp = r->bp.n / 7;
    double tK = r->tK, beta = s->mc_kT > 0.0 ? 1.0 / s->mc_kT : 1.0 / (BOLTZ_KCAL_MOL * tK);
    double (*fbp)[6][3] = calloc(nbp + 1, sizeof(double[6][3]));
    int *nt_excess = malloc(sizeof(int) * s->nmp);
    for (int i = 0; i < s->nmp; i++) nt_excess[i] = -s->p.max_bp_per_nt;
    for (int64_t b = 0; b < nbp; b++) { r->bp_status[b] = 0; r->ene_bp[b] = 0.0; }
    for (int64_t b = 0; b < nbp; b++) {
        if (!bp_mc && s->p.nstep_bp_mc > 0 && !r->bp_status_mc[b]) continue;
        const int32_t *mp = r->bp.a + 7 * b;
        double dG = bp_prefactor(s, r, b, tK);
        if (s->p.bp_model == 5 && dG >= 0.0) continue;
        double u, fb[6][3];
        if (!bp_geom(s, x, mp, &u, fb)) continue;
        double e = s->bpp[mp[6]].U0 * dG * exp(-u);
        if (e <= s->p.bp_cutoff_energy) {
            r->bp_status[b] = 1;
            r->ene_bp[b] = e;
            for (int q = 0; q < 6; q++) for (int c = 0; c < 3; c++) fbp[b][q][c] = e * fb[q][c];
            nt_excess[mp[0] - 1] += 1;
            nt_excess[mp[1] - 1] += 1;
        }
    }
    if (bp_mc) bp_limit_mc(s, irep, nt_excess, beta);
    for (int64_t b = 0; b < nbp; b++) {
        if (!r->bp_status[b]) continue;
        const int32_t *mp = r->bp.a + 7 * b;
        for (int q = 0; q < 6; q++) for (int c = 0; c < 3; c++) F3(f, mp[q] - 1)[c] += fbp[b][q][c];
    }
    s->flg_bp_energy = 1;
    free(fbp);
    free(nt_excess);
}

/* force (src/force.F90:27-69), single thread */
/* centre of mass and Rg (src/force_rg.F90:17-27, src/energy_rg.F90:17-27) */
static double rg_of(const orc_sys *s, const double *x, double *cm) {
    int n = s->nmp;
    for (int k = 0; k < 3; k++) {
        double t = 0.0;
        for (int i = 0; i < n; i++) t += x[3 * i + k];
        cm[k] = t / n;
    }
    double s2 = 0.0;
    for (int i = 0; i < n; i++) {
        double v[3] = {x[3 * i] - cm[0], x[3 * i + 1] - cm[1], x[3 * i + 2] - cm[2]};
        s2 = s2 + dot3(v, v);
    }
    return sqrt(s2 / n);
}

/* external terms after the pair and bonded sum, in the reference's order
 * (src/force.F90:62-75): stage inside the sum, then tweezers, Rg bias and
 * restraints added to the forces */
static void force_external(orc_sys *s, int irep, const double *x, double *f) {
    const sis_external *e = &s->x;
    if (e->stage) {   /* src/force_stage.F90:17-30 */
        double coef = 24.0 * e->stage_eps / e->stage_sigma;
        for (int i = 0; i < s->nmp; i++) {
            double z = x[3 * i + 2];
            if (z < 2.220446049250313e-16) z = 2.220446049250313e-16;
            z = e->stage_sigma / z;
            f[3 * i + 2] = f[3 * i + 2] + coef * (2 * pow(z, 13) - pow(z, 7));
        }
    }
    for (int k = 0; k < e->n_dcf; k++) {   /* src/force_tweezers.F90:30-36 */
        int i = s->x_dcf_pairs[2 * k] - 1, j = s->x_dcf_pairs[2 * k + 1] - 1;
        const double *fv = s->x_dcf_f + 3 * ((size_t)irep * e->n_dcf + k);
        for (int c = 0; c < 3; c++) { f[3 * i + c] += fv[c]; f[3 * j + c] -= fv[c]; }
    }
    for (int k = 0; k < e->n_fr; k++) {   /* src/force_tweezers.F90:38-48 */
        for (int end = 0; end < 2; end++) {
            int i = s->x_fr_pairs[2 * k + end] - 1;
            double p[3], v[3];
            for (int c = 0; c < 3; c++) p[c] = s->x_fr_init[6 * k + 3 * end + c] + s->x_fr_velo[6 * k + 3 * end + c] * (double)s->istep;
            pbc_vec_d(s, p, F3(x, i), v);
            for (int c = 0; c < 3; c++) f[3 * i + c] = f[3 * i + c] + s->x_fr_k[2 * k + end] * v[c];
        }
    }
    if (e->rg) {   /* src/force_rg.F90:17-39 */
        double cm[3], rg = rg_of(s, x, cm), r0 = s->x_rg0[irep];
        if (!(e->rg == 2 && rg <= r0)) {
            double factor = -e->rg_k * (rg - r0) / (rg * s->nmp);
            for (int i = 0; i < s->nmp; i++)
                for (int c = 0; c < 3; c++) f[3 * i + c] = f[3 * i + c] + factor * (x[3 * i + c] - cm[c]);
        }
    }
    for (int k = 0; k < e->n_rest; k++) {   /* src/force_restraint.F90:18-40 */
        int i = s->x_rest_id[3 * k] - 1, j = s->x_rest_id[3 * k + 1] - 1, ty = s->x_rest_id[3 * k + 2];
        double v[3];
        pbc_vec_d(s, F3(x, i), F3(x, j), v);
        double r = sqrt(dot3(v, v));
        if (r >= s->x_rest_rcut[k]) continue;
        double eps = s->x_rest_para[3 * k], d = s->x_rest_para[3 * k + 1], sg = s->x_rest_para[3 * k + 2];
        double t = tanh((r - d) / sg);
        double coef = 0.5 * eps / sg * (1.0 - t * t) / r;
        for (int c = 0; c < 3; c++) f[3 * i + c] = f[3 * i + c] - coef * v[c];
        if (ty == 1) for (int c = 0; c < 3; c++) f[3 * j + c] = f[3 * j + c] + coef * v[c];
    }
}

/* energies of the external terms into E[7..10] (src/energy_stage.F90,
 * energy_tweezers.F90, energy_rg.F90, energy_restraint.F90); returns -1
 * where energy_stage stops the run (a bead below 0.85 stage_sigma) */
static int energy_external(orc_sys *s, int irep, const double *x, double *E) {
    const sis_external *e = &s->x;
    if (e->stage) {
        double es = 0.0;
        for (int i = 0; i < s->nmp; i++) if (x[3 * i + 2] < 0.85 * e->stage_sigma) return -1;
        for (int i = 0; i < s->nmp; i++) {
            double z = e->stage_sigma / x[3 * i + 2];
            es = es + 4 * e->stage_eps * (pow(z, 12) - pow(z, 6));
        }
        E[7] = es;
    }
    if (e->n_dcf > 0 || e->n_fr > 0) {
        double et = 0.0;
        for (int k = 0; k < e->n_dcf; k++) {
            int i = s->x_dcf_pairs[2 * k] - 1, j = s->x_dcf_pairs[2 * k + 1] - 1;
            double v[3];
            pbc_vec_d(s, F3(x, i), F3(x, j), v);
            et = et - dot3(v, s->x_dcf_f + 3 * ((size_t)irep * e->n_dcf + k));
        }
        E[8] = et;
    }
    if (e->rg) {
        double cm[3], rg = rg_of(s, x, cm), r0 = s->x_rg0[irep];
        E[9] = (e->rg == 2 && rg <= r0) ? 0.0 : 0.5 * e->rg_k * (rg - r0) * (rg - r0);
    }
    for (int k = 0; k < e->n_rest; k++) {
        int i = s->x_rest_id[3 * k] - 1, j = s->x_rest_id[3 * k + 1] - 1;
        double v[3];
        pbc_vec_d(s, F3(x, i), F3(x, j), v);
        double r = sqrt(dot3(v, v));
        if (r >= s->x_rest_rcut[k]) continue;
        double eps = s->x_rest_para[3 * k], d = s->x_rest_para[3 * k + 1], sg = s->x_rest_para[3 * k + 2];
        E[10] = E[10] - 0.5 * eps * (1.0 - tanh((r - d) / sg));
    }
    return 0;
}

void orc_force(orc_sys *s, int32_t irep1, int32_t bp_mc, double *forces) {
    int irep = irep1 - 1;
    if (s->p.bp_model != 5) bp_mc = 0;   /* models 1/3 run without the limit MC */
    const double *x = s->xyz + (size_t)irep * s->nmp * 3;
    memset(forces, 0, sizeof(double) * 3 * s->nmp);
    force_bond(s, x, forces);
    force_angl(s, x, forces);
    force_dih(s, x, forces);
    force_bp(s, irep, bp_mc, x, forces);
    force_wca(s, irep, x, forces);
    if (s->p.ele_on) force_ele(s, irep, x, forces);
    if (s->x_on) force_external(s, irep, x, forces);
}

static void *dup_n(const void *src, size_t bytes) {
    if (!src || bytes == 0) return NULL;
    void *p = malloc(bytes);
    memcpy(p, src, bytes);
    return p;
}

static void free_external(orc_sys *s) {
    free(s->x_dcf_pairs); free(s->x_fr_pairs); free(s->x_rest_id); free(s->x_rg0); free(s->x_dcf_f);
    free(s->x_fr_k); free(s->x_fr_init); free(s->x_fr_velo); free(s->x_rest_rcut); free(s->x_rest_para);
    s->x_dcf_pairs = s->x_fr_pairs = s->x_rest_id = NULL;
    s->x_rg0 = s->x_dcf_f = s->x_fr_k = s->x_fr_init = s->x_fr_velo = s->x_rest_rcut = s->x_rest_para = NULL;
    s->x_on = 0;
}

void orc_set_step(orc_sys *s, int64_t istep) { s->istep = istep; }
void orc_get_box(orc_sys *s, double *box) { for (int k = 0; k < 3; k++) box[k] = s->p.box[k]; }

/* The pairs of a CT/BPSEQ/BPL file (1-based bead ids, either order) as
 * read_ss keeps them (src/read_ss.F90:126-205): ends of a chain, pairs
 * closer than the minimum loop, isolated pairs (models 3-5) and
 * non-canonical pairs are dropped.  npair < 0 clears the restriction.
 * Returns the number of pairs kept. */
static void rg_schedule_apply(orc_sys *s) {
    s->x.rg_k = s->rgs_k[s->irgs - 1];
    for (int r = 0; r < s->nrep; r++) s->x_rg0[r] = s->rgs_0[s->irgs - 1];
    for (int k = 1; k <= s->nrep; k++) s->lab2rg[k] = s->rgs_0[s->irgs - 1];
    s->rgs_next = s->irgs < s->rgs_n ? s->rgs_step[s->irgs] : 0;
}

/* Timed Rg bias (read_timed_bias_rg; job_md's set-up and its switch at the
 * end of step next - 1).  Needs the Rg bias of orc_set_external. */
void orc_set_rg_schedule(orc_sys *s, int32_t n, const int64_t *steps, const double *k, const double *r0) {
    free(s->rgs_step); free(s->rgs_k); free(s->rgs_0);
    s->rgs_step = NULL; s->rgs_k = NULL; s->rgs_0 = NULL; s->rgs_n = 0; s->rgs_next = 0;
    if (n <= 0 || !s->x_on || !s->x.rg) return;
    s->rgs_n = n;
    s->rgs_step = malloc(sizeof(int64_t) * n); memcpy(s->rgs_step, steps, sizeof(int64_t) * n);
    s->rgs_k = malloc(sizeof(double) * n); memcpy(s->rgs_k, k, sizeof(double) * n);
    s->rgs_0 = malloc(sizeof(double) * n); memcpy(s->rgs_0, r0, sizeof(double) * n);
    int i = 1;
    if (s->istep >= steps[n - 1]) i = n;
    else while (i < n && steps[i] <= s->istep) i++;
    s->irgs = i;
    rg_schedule_apply(s);
}

/* [PBC_box] resizing: pbc_resize (src/pbc.F90:79-138) every resize_step
 * steps; target may be NULL */
void orc_set_box_resize(orc_sys *s, int64_t resize_step, const double *change, const double *target) {
    s->rz_on = resize_step > 0;
    s->rz_step = resize_step;
    s->rz_has_target = target != NULL;
    for (int k = 0; k < 3; k++) {
        s->rz_change[k] = change ? change[k] : 0.0;
        s->rz_target[k] = target ? target[k] : 0.0;
        s->rz_orig[k] = s->p.box[k];
        s->rz_count[k] = 0;
    }
}

static void pbc_resize(orc_sys *s) {
    const double SMALL = 2.220446049250313e-16;
    int reach[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) {
        if (fabs(s->rz_change[i]) < SMALL) { reach[i] = 1; continue; }
        if (s->rz_has_target && fabs(s->p.box[i] - s->rz_target[i]) < SMALL) { reach[i] = 1; continue; }
        s->rz_count[i]++;
        double x = s->rz_orig[i] + s->rz_change[i] * s->rz_count[i];
        if (s->rz_change[i] < 0.0) {
            if (s->rz_has_target && x - s->rz_target[i] < SMALL) { s->p.box[i] = s->rz_target[i]; reach[i] = 1; }
            else s->p.box[i] = x;
        } else {
            if (s->rz_has_target && x - s->rz_target[i] > SMALL) { s->p.box[i] = s->rz_target[i]; reach[i] = 1; }
            else s->p.box[i] = x;
        }
    }
    if (reach[0] && reach[1] && reach[2]) s->rz_on = 0;
}

/* Annealing (src/main.F90:90-97, src/job_md.F90:207-230,600-612): the
 * electrostatics are set once at the first temperature, the force's MC
 * keeps the input kT, and the replicas' temperature follows the schedule. */
void orc_set_anneal(orc_sys *s, int32_t n, const int64_t *steps, const double *temps) {
    free(s->an_step); free(s->an_T);
    s->an_step = NULL; s->an_T = NULL; s->an_n = 0; s->an_next = 0; s->mc_kT = 0.0;
    if (n <= 0) return;
    s->an_n = n;
    s->an_step = malloc(sizeof(int64_t) * n); memcpy(s->an_step, steps, sizeof(int64_t) * n);
    s->an_T = malloc(sizeof(double) * n); memcpy(s->an_T, temps, sizeof(double) * n);
    s->ianneal = s->istep >= steps[n - 1] ? n : 1;
    s->an_next = s->ianneal < n ? steps[s->ianneal] : 0;
    for (int r = 0; r < s->nrep; r++) {
        init_ele(s, r, temps[0], s->rep[r].ionic);
        s->rep[r].tK = temps[s->ianneal - 1];
    }
    s->mc_kT = BOLTZ_KCAL_MOL * s->p.tempK;
}

int32_t orc_set_ss(orc_sys *s, int32_t npair, const int32_t *pairs) {
    free(s->ss_map);
    s->ss_map = NULL;
    if (npair < 0) return 0;
    int n = s->nmp, kept = 0;
    s->ss_map = calloc((size_t)n * n, 1);
    for (int k = 0; k < npair; k++) {
        int imp = pairs[2 * k], jmp = pairs[2 * k + 1];
        if (jmp == 0) continue;
        if (jmp < imp) { int t = imp; imp = jmp; jmp = t; }
        imp--; jmp--;
        int ich = s->chain_of[imp], jch = s->chain_of[jmp];
        int i = s->lmp[imp], j = s->lmp[jmp];
        if (i == 1 || i == s->nmp_chain[ich]) continue;
        if (j == 1 || j == s->nmp_chain[jch]) continue;
        if (ich == jch && i + s->p.bp_min_loop >= j) continue;
        const int *si = s->seq + s->chain_start[ich] - 1, *sj = s->seq + s->chain_start[jch] - 1;
        int m = s->p.bp_model;
        if ((m == 3 || m == 4 || m == 5) && !is_complement(si[i - 1], sj[j + 1]) && !is_complement(si[i + 1], sj[j - 1]))
            continue;
        int a = si[i], b = sj[j];
        int canon = (a == G_ && b == C_) || (a == C_ && b == G_) || (a == A_ && b == U_) || (a == U_ && b == A_) ||
                    (a == G_ && b == U_) || (a == U_ && b == G_);
        if (!canon) continue;
        s->ss_map[(size_t)imp * n + jmp] = s->ss_map[(size_t)jmp * n + imp] = 1;
        kept++;
    }
    return kept;
}

void orc_set_external(orc_sys *s, const sis_external *e) {
    free_external(s);
    if (!e) return;
    s->x = *e;
    s->x_on = 1;
    s->x_dcf_pairs = dup_n(e->dcf_pairs, sizeof(int32_t) * 2 * e->n_dcf);
    s->x_dcf_f = dup_n(e->dcf_forces, sizeof(double) * 3 * e->n_dcf * s->nrep);
    s->x_fr_pairs = dup_n(e->fr_pairs, sizeof(int32_t) * 2 * e->n_fr);
    s->x_fr_k = dup_n(e->fr_k, sizeof(double) * 2 * e->n_fr);
    s->x_fr_init = dup_n(e->fr_init, sizeof(double) * 6 * e->n_fr);
    s->x_fr_velo = dup_n(e->fr_velo, sizeof(double) * 6 * e->n_fr);
    s->x_rg0 = dup_n(e->rg_0, sizeof(double) * s->nrep);
    s->x_rest_id = dup_n(e->rest_id, sizeof(int32_t) * 3 * e->n_rest);
    s->x_rest_rcut = dup_n(e->rest_rcut, sizeof(double) * e->n_rest);
    s->x_rest_para = dup_n(e->rest_para, sizeof(double) * 3 * e->n_rest);
    s->rg_remd = 0;
    s->dcf_remd = 0;
}

static void dcf_set_forces(orc_sys *s, int r, double k) {
    for (int c = 0; c < 3; c++) s->x_dcf_f[3 * r + c] = k * s->dcf_dir[c];
}

/* f-REMD over the one constant-force pair (src/init_tweezers.F90:24-40):
 * force = k[label] x normalised direction */
void orc_set_rep_dcf(orc_sys *s, const double *k, const double *dir) {
    if (!s->x_on || s->x.n_dcf != 1) return;
    double n = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (int c = 0; c < 3; c++) s->dcf_dir[c] = dir[c] / n;
    for (int r = 0; r < s->nrep; r++) { s->lab2dcf[s->rep2lab[r + 1]] = k[r]; dcf_set_forces(s, r, k[r]); }
    s->dcf_remd = 1;
}

/* REMD over the Rg-bias target (REPT%RG): rg0 of each replica's label
 * (init_bias_rg, src/init_bias_rg.F90:13-16); needs the Rg bias */
void orc_set_rep_rg(orc_sys *s, const double *rg0) {
    if (!s->x_on || !s->x.rg) return;
    for (int r = 0; r < s->nrep; r++) { s->lab2rg[s->rep2lab[r + 1]] = rg0[r]; s->x_rg0[r] = rg0[r]; }
    s->rg_remd = 1;
}

/* ---------------- energies (src/energy_*.F90) ---------------- */
static double mp_angle(orc_sys *s, const double *x, int a, int b, int c) {   /* energy_bp_limit_triplet.F90:199-218 */
    double v12[3], v32[3];
    pbc_vec_d(s, F3(x, a), F3(x, b), v12);
    pbc_vec_d(s, F3(x, c), F3(x, b), v32);
    double co = dot3(v32, v12) / sqrt(dot3(v12, v12) * dot3(v32, v32));
    if (co > 1.0) co = 1.0; else if (co < -1.0) co = -1.0;
    return acos(co);
}
static double mp_dihedral(orc_sys *s, const double *x, int a, int b, int c, int d) {   /* :220-240 */
    double v12[3], v32[3], v34[3], m[3], n[3];
    pbc_vec_d(s, F3(x, a), F3(x, b), v12);
    pbc_vec_d(s, F3(x, c), F3(x, b), v32);
    pbc_vec_d(s, F3(x, c), F3(x, d), v34);
    cross(v12, v32, m);
    cross(v32, v34, n);
    return atan2(dot3(v12, n) * sqrt(dot3(v32, v32)), dot3(m, n));
}
static double energy_bp(orc_sys *s, int irep, double tK, int bp_mc, const double *x) {
    /* src/energy_bp_limit_triplet.F90:35-195 and src/energy_bp_triplet.F90 */
    rep_t *r = &s->rep[irep];
    int64_t nbp = r->bp.n / 7;
    if (!s->flg_bp_energy) {
        double beta = 1.0 / (BOLTZ_KCAL_MOL * tK);
        int *nt_excess = malloc(sizeof(int) * s->nmp);
        for (int i = 0; i < s->nmp; i++) nt_excess[i] = -s->p.max_bp_per_nt;
        for (int64_t b = 0; b < nbp; b++) { r->bp_status[b] = 0; r->ene_bp[b] = 0.0; }
        for (int64_t b = 0; b < nbp; b++) {
            if (!bp_mc && s->p.nstep_bp_mc > 0 && !r->bp_status_mc[b]) continue;
            const int32_t *mp = r->bp.a + 7 * b;
            const sis_bp_para *bpp = &s->bpp[mp[6]];
            double dG = bp_prefactor(s, r, b, tK);
            if (s->p.bp_model == 5 && dG >= 0.0) continue;
            double v[3];
            pbc_vec_d(s, F3(x, mp[0] - 1), F3(x, mp[1] - 1), v);
            double d = sqrt(dot3(v, v)) - bpp->bond_r;
            if (fabs(d) > bpp->cutoff_ddist) continue;
            double u = bpp->bond_k * (d * d), th, ph;
            th = mp_angle(s, x, mp[0] - 1, mp[1] - 1, mp[3] - 1); u = u + bpp->angl_k1 * ((th - bpp->angl_theta1) * (th - bpp->angl_theta1));
            th = mp_angle(s, x, mp[2] - 1, mp[0] - 1, mp[1] - 1); u = u + bpp->angl_k2 * ((th - bpp->angl_theta2) * (th - bpp->angl_theta2));
            th = mp_angle(s, x, mp[0] - 1, mp[1] - 1, mp[5] - 1); u = u + bpp->angl_k3 * ((th - bpp->angl_theta3) * (th - bpp->angl_theta3));
            th = mp_angle(s, x, mp[4] - 1, mp[0] - 1, mp[1] - 1); u = u + bpp->angl_k4 * ((th - bpp->angl_theta4) * (th - bpp->angl_theta4));
            ph = mp_dihedral(s, x, mp[2] - 1, mp[0] - 1, mp[1] - 1, mp[3] - 1); u = u + bpp->dihd_k1 * (1.0 + cos(ph + bpp->dihd_phi1));
            ph = mp_dihedral(s, x, mp[4] - 1, mp[0] - 1, mp[1] - 1, mp[5] - 1); u = u + bpp->dihd_k2 * (1.0 + cos(ph + bpp->dihd_phi2));
            u = bpp->U0 * dG * exp(-u);
            if (u <= s->p.bp_cutoff_energy) {
                r->ene_bp[b] = u;
                r->bp_status[b] = 1;
                nt_excess[mp[0] - 1] += 1;
                nt_excess[mp[1] - 1] += 1;
            }
        }
        if (bp_mc) bp_limit_mc(s, irep, nt_excess, beta);
        free(nt_excess);
    }
    /* energy_bp_limit_triplet sums formed pairs only
     * (src/energy_bp_limit_triplet.F90:195); energy_bp_triplet sums ene_bp
     * of every candidate (src/energy_bp_triplet.F90:88), which still holds
     * the energies of pairs removed by an earlier limit MC */
    double e = 0.0;
    for (int64_t b = 0; b < nbp; b++) if (r->bp_status[b] || !bp_mc) e += r->ene_bp[b];
    return e;
}

void orc_energy(orc_sys *s, int32_t irep1, double tK, int32_t bp_mc, int32_t bp_energy_valid, double *E) {
    int irep = irep1 - 1;
    if (s->p.bp_model != 5) bp_mc = 0;
    rep_t *r = &s->rep[irep];
    const double *x = s->xyz + (size_t)irep * s->nmp * 3;
    for (int k = 0; k < 11; k++) E[k] = 0.0;
    s->flg_bp_energy = bp_energy_valid;
    /* bond (src/energy_bond.F90:16-24) */
    for (int c = 0; c < s->nchains; c++) {
        int n = s->nmp_chain[c], st = s->chain_start[c];
        int nb = n - 1 + (s->moltype[c] == 2);
        for (int b = 0; b < nb; b++) {
            int i1 = (b < n - 1) ? st + b : st + n - 1, i2 = (b < n - 1) ? st + b + 1 : st;
            double v[3];
            pbc_vec_d(s, F3(x, i1), F3(x, i2), v);
            double d = sqrt(dot3(v, v));
            E[1] = E[1] + 0.5 * s->p.bond_k * ((d - s->p.bond_r0) * (d - s->p.bond_r0));
        }
    }
    /* angle (src/energy_angl_ReB.F90:17-30, src/energy_angl.F90) */
    {
        int (*al)[3] = malloc(sizeof(int[3]) * (s->nmp + 2 * s->nchains));
        int na = angle_list(s, al);
        double co0 = cos(s->p.angl_t0);
        for (int a = 0; a < na; a++) {
            if (s->p.angl_reb) {
                double v12[3], v32[3];
                pbc_vec_d(s, F3(x, al[a][0]), F3(x, al[a][1]), v12);
                pbc_vec_d(s, F3(x, al[a][2]), F3(x, al[a][1]), v32);
                double co = dot3(v32, v12) / sqrt(dot3(v12, v12) * dot3(v32, v32));
                E[2] = E[2] + 0.5 * s->p.angl_k * ((co - co0) * (co - co0)) / (1.0 - co * co);
            } else {
                double th = mp_angle(s, x, al[a][0], al[a][1], al[a][2]);
                E[2] = E[2] + 0.5 * s->p.angl_k * ((th - s->p.angl_t0) * (th - s->p.angl_t0));
            }
        }
        free(al);
    }
    /* dihedral (src/energy_dih_exp.F90:18-40, src/energy_dih_cos.F90) */
    if (s->p.dih_type) {
        int (*dl)[4] = malloc(sizeof(int[4]) * (s->nmp + 3 * s->nchains));
        int nd = dih_list(s, dl);
        double coef = -0.5 * s->p.dih_w;
        for (int a = 0; a < nd; a++) {
            double v12[3], v23[3], v34[3], c1[3], c2[3];
            pbc_vec_d(s, F3(x, dl[a][1]), F3(x, dl[a][0]), v12);
            pbc_vec_d(s, F3(x, dl[a][2]), F3(x, dl[a][1]), v23);
            pbc_vec_d(s, F3(x, dl[a][3]), F3(x, dl[a][2]), v34);
            cross(v12, v23, c1);
            cross(v23, v34, c2);
            double dih = atan2(sqrt(dot3(v23, v23)) * dot3(v12, c2), dot3(c1, c2));
            if (s->p.dih_type == 2) E[3] = E[3] - s->p.dih_k * exp(coef * ((dih - s->p.dih_p0) * (dih - s->p.dih_p0)));
            else E[3] = E[3] + s->p.dih_k * (1.0 + cos(dih + s->p.dih_p0));
        }
        free(dl);
    }
    E[4] = energy_bp(s, irep, tK, bp_mc, x);
    /* WCA (src/energy_wca.F90:19-33) */
    {
        double e = 0.0;
        for (int64_t k = 0; k < r->wca.n / 2; k++) {
            double v[3];
            pbc_vec_d(s, F3(x, r->wca.a[2 * k] - 1), F3(x, r->wca.a[2 * k + 1] - 1), v);
            double d = sqrt(dot3(v, v));
            if (d >= s->p.wca_sigma) continue;
            d = s->p.wca_sigma / d;
            double d6 = d * d * d * d * d * d;
            e = e + s->p.wca_eps * (d6 * d6 - 2 * d6 + 1.0);
        }
        E[5] = e;
    }
    /* DH (src/energy_ele_DH.F90:19-42) */
    if (s->p.ele_on) {
        double e = 0.0, rc = 1.0 / r->lambdaD;
        for (int64_t k = 0; k < r->ele.n / 2; k++) {
            double v[3];
            pbc_vec_d(s, F3(x, r->ele.a[2 * k] - 1), F3(x, r->ele.a[2 * k + 1] - 1), v);
            double dist = sqrt(dot3(v, v));
            if (dist > r->ele_cutoff) continue;
            e = e + r->ele_coef / dist * exp(-dist * rc);
        }
        E[6] = e;
    }
    if (s->x_on && energy_external(s, irep, x, E) < 0) {
        for (int k = 0; k < 11; k++) E[k] = NAN;
        return;
    }
    double tot = 0.0;
    for (int k = 1; k <= 10; k++) tot += E[k];
    E[0] = tot;
}

void orc_bp_state(orc_sys *s, int32_t irep, uint8_t *status, double *ene) {
    rep_t *r = &s->rep[irep - 1];
    int64_t nbp = r->bp.n / 7;
    memcpy(status, r->bp_status, nbp);
    memcpy(ene, r->ene_bp, nbp * sizeof(double));
}

/* ---------------- set up ---------------- */
static double md_mass(int sq) {   /* src/job_md.F90:116-133 */
    switch (sq) { case A_: return 328.212; case G_: return 344.212; case C_: return 304.182;
                  case U_: return 305.164; default: return 320.443; }
}

orc_sys *orc_create(const sis_topo *topo, const sis_params *prm, int32_t nrep, const double *temps) {
    JOUL2KCAL_MOL = (1.0 / KCAL2JOUL) * N_AVO;
    BOLTZ_KCAL_MOL = BOLTZ_J * JOUL2KCAL_MOL;
    orc_sys *s = calloc(1, sizeof *s);
    s->p = *prm;
    s->nmp = topo->nmp; s->nchains = topo->nchains; s->nrep = nrep;
    s->nmp_chain = malloc(sizeof(int) * s->nchains);
    s->moltype = malloc(sizeof(int) * s->nchains);
    s->chain_start = malloc(sizeof(int) * s->nchains);
    s->seq = malloc(sizeof(int) * s->nmp);
    s->chain_of = malloc(sizeof(int) * s->nmp);
    s->lmp = malloc(sizeof(int) * s->nmp);
    s->has_charge = malloc(s->nmp);
    int imp = 0;
    for (int c = 0; c < s->nchains; c++) {
        s->nmp_chain[c] = topo->nmp_chain[c];
        s->moltype[c] = topo->moltype[c];
        s->chain_start[c] = imp;
        for (int i = 1; i <= s->nmp_chain[c]; i++, imp++) { s->chain_of[imp] = c; s->lmp[imp] = i; }
    }
    for (int i = 0; i < s->nmp; i++) {
        s->seq[i] = topo->seq[i];
        s->has_charge[i] = prm->dummy_has_charge ? 1 : (s->seq[i] != D_);   /* src/init_ele.F90:58-68 */
    }
    for (int k = 0; k < topo->n_no_charge; k++) s->has_charge[topo->no_charge[k] - 1] = 0;
    /* base-pair parameter blocks and their mirrors (src/read_force_field.F90:203-207) */
    for (int t = 0; t < 3; t++) {
        sis_bp_para b = prm->bp[t], m = b;
        m.angl_k1 = b.angl_k2; m.angl_k2 = b.angl_k1; m.angl_k3 = b.angl_k4; m.angl_k4 = b.angl_k3;
        s->bpp[t + 1] = b; s->bpp[t + 4] = m;
    }
    init_bp_tables(s);
    size_t n3 = (size_t)nrep * s->nmp * 3;
    s->xyz = calloc(n3, sizeof(double)); s->velos = calloc(n3, sizeof(double));
    s->accels = calloc(n3, sizeof(double)); s->move = calloc(n3, sizeof(double));
    s->rep = calloc(nrep, sizeof(rep_t));
    /* friction and masses (src/job_md.F90:94-133): the viscosity in
     * internal units differs between the Langevin and the Brownian
     * integrator (lines 98-100 and 106-108) */
    s->mass = malloc(sizeof(double) * s->nmp);
    s->fric = malloc(sizeof(double) * s->nmp);
    double v = prm->integrator == 2 ? (JOUL2KCAL_MOL * 1.0e-18) * prm->viscosity_pas
                                    : prm->viscosity_pas * sqrt(1.0e3 / KCAL2JOUL) * N_AVO * 1.0e-20;
    for (int i = 0; i < s->nmp; i++) { s->mass[i] = md_mass(s->seq[i]); s->fric[i] = 6.0 * PI_ * v * 10.0; }
    /* replicas: labels, values and the odd/even pair table of a
     * one-dimensional temperature ladder (src/init_replica.F90:70-125,
     * src/var_replica.F90:162-259) */
    s->rep2lab = malloc(sizeof(int) * (nrep + 1));
    s->lab2rep = malloc(sizeof(int) * (nrep + 1));
    s->lab2val = malloc(sizeof(double) * (nrep + 1));
    s->lab2ion = malloc(sizeof(double) * (nrep + 1));
    s->lab2rg = calloc(nrep + 1, sizeof(double));
    s->lab2dcf = calloc(nrep + 1, sizeof(double));
    for (int k = 0; k <= nrep; k++) s->lab2ion[k] = prm->ionic_strength;
    for (int k = 1; k <= nrep; k++) { s->rep2lab[k] = k; s->lab2rep[k] = k; s->lab2val[k] = temps ? temps[k - 1] : prm->tempK; }
    s->pair_tb = NULL;
    { int32_t one = nrep; orc_set_rep_dims(s, 1, &one); }
    for (int r = 0; r < nrep; r++) {
        s->rep[r].tK = (nrep > 1) ? s->lab2val[s->rep2lab[r + 1]] : prm->tempK;
        s->rep[r].ionic = prm->ionic_strength;
        init_ele(s, r, s->rep[r].tK, s->rep[r].ionic);
    }
    return s;
}

void orc_set_temps(orc_sys *s, const double *temps) {
    for (int r = 0; r < s->nrep; r++) { s->rep[r].tK = temps[r]; init_ele(s, r, temps[r], s->rep[r].ionic); }
}

/* REMD over the ionic strength (REPT%ION): label k carries ionic[k-1];
 * each replica takes its label's (src/init_ele.F90, energy_replica and
 * replica_exchange treat it like the temperature) */
/* make_replica_exchange_pair_tb (src/var_replica.F90:162-259) for ndim
 * replica dimensions of nv[d] values each (labels laid out as
 * src/init_replica.F90:66-93 does, the first dimension slowest).  Table
 * icounter = (iparity - 1) * ndim + idimn pairs labels n_continue apart
 * along dimension idimn, odd pairs first. */
void orc_set_rep_dims(orc_sys *s, int32_t ndim, const int32_t *nv) {
    int R = s->nrep;
    free(s->pair_tb);
    s->pair_tb = calloc((size_t)2 * ndim * (R + 2), sizeof(int));
    s->pair_sets = 2 * ndim;
    int icounter = 0;
    for (int iparity = 1; iparity <= 2; iparity++) {
        int n_division = 1;
        for (int idimn = 1; idimn <= ndim; idimn++) {
            icounter++;
            int *tb = s->pair_tb + (size_t)(icounter - 1) * (R + 2);
            int nrep_v = nv[idimn - 1];
            int n_division_pre = n_division;
            n_division *= nrep_v;
            int n_continue = R / n_division;
            int iset = 0;
            for (int icycle = 1; icycle <= n_division_pre; icycle++) {
                int post_zero = 0, start;
                if (iparity == 1) {
                    if (nrep_v % 2 == 1) post_zero = 1;
                    start = 1;
                } else {
                    for (int ic = 1; ic <= n_continue; ic++) tb[++iset] = 0;
                    if (nrep_v % 2 == 0) post_zero = 1;
                    start = 2;
                }
                for (int irep = start; irep <= nrep_v - 1; irep += 2)
                    for (int ic = 1; ic <= 2 * n_continue; ic++) {
                        iset++;
                        tb[iset] = (ic <= n_continue) ? iset + n_continue : iset - n_continue;
                    }
                if (post_zero)
                    for (int ic = 1; ic <= n_continue; ic++) tb[++iset] = 0;
            }
        }
    }
    s->idx_pair = 0;
}

void orc_pair_table(orc_sys *s, int32_t *out) {
    memcpy(out, s->pair_tb, sizeof(int) * (size_t)s->pair_sets * (s->nrep + 2));
}

void orc_set_ionic(orc_sys *s, const double *ionic) {
    for (int k = 1; k <= s->nrep; k++) s->lab2ion[k] = ionic[k - 1];
    for (int r = 0; r < s->nrep; r++) {
        s->rep[r].ionic = s->lab2ion[s->rep2lab[r + 1]];
        init_ele(s, r, s->rep[r].tK, s->rep[r].ionic);
    }
}

void orc_destroy(orc_sys *s) {
    if (!s) return;
    free_external(s);
    free(s->ss_map);
    free(s->an_step); free(s->an_T);
    free(s->rgs_step); free(s->rgs_k); free(s->rgs_0);
    for (int r = 0; r < s->nrep; r++) {
        free(s->rep[r].wca.a); free(s->rep[r].ele.a); free(s->rep[r].bp.a); free(s->rep[r].bpcoef.a);
        free(s->rep[r].bp_status); free(s->rep[r].bp_status_mc); free(s->rep[r].ene_bp);
    }
    free(s->rep); free(s->nmp_chain); free(s->moltype); free(s->chain_start); free(s->seq);
    free(s->chain_of); free(s->lmp); free(s->has_charge); free(s->xyz); free(s->velos);
    free(s->accels); free(s->move); free(s->mass); free(s->fric); free(s->rep2lab);
    free(s->lab2rep); free(s->lab2val); free(s->lab2ion); free(s->lab2rg); free(s->lab2dcf); free(s->pair_tb);
    free(s);
}

void orc_set_xyz(orc_sys *s, const double *xyz) { memcpy(s->xyz, xyz, sizeof(double) * 3 * s->nmp * s->nrep); }
void orc_set_state(orc_sys *s, const double *xyz, const double *velos, const double *accels) {
    size_t n = sizeof(double) * 3 * s->nmp * s->nrep;
    if (xyz) memcpy(s->xyz, xyz, n);
    if (velos) memcpy(s->velos, velos, n);
    if (accels) memcpy(s->accels, accels, n);
}
void orc_get_state(orc_sys *s, double *xyz, double *velos, double *accels) {
    size_t n = sizeof(double) * 3 * s->nmp * s->nrep;
    if (xyz) memcpy(xyz, s->xyz, n);
    if (velos) memcpy(velos, s->velos, n);
    if (accels) memcpy(accels, s->accels, n);
}
void orc_tables(orc_sys *s, double *dh, double *ds, double *cut) {
    memcpy(dh, s->bp3_dH + 1, 468 * sizeof(double));
    memcpy(ds, s->bp3_dS + 1, 468 * sizeof(double));
    cut[0] = s->bp_cutoff_dist;
    for (int t = 1; t <= 6; t++) cut[t] = s->bpp[t].cutoff_ddist;
}
void orc_ele_params(orc_sys *s, int32_t irep, double *out) {
    rep_t *r = &s->rep[irep - 1];
    out[0] = r->lambdaD; out[1] = r->ele_coef; out[2] = r->ele_cutoff; out[3] = r->diele;
}
void orc_set_tape(orc_sys *s, int32_t irep, const uint32_t *words, int64_t n) {
    tape_t *t = irep < 0 ? &s->tape_rep : &s->rep[irep - 1].tape;
    t->w = words; t->n = n; t->pos = 0;
}
int64_t orc_tape_pos(orc_sys *s, int32_t irep) { return irep < 0 ? s->tape_rep.pos : s->rep[irep - 1].tape.pos; }
void orc_boxmuller(orc_sys *s, int32_t irep, int64_t n, double *out) {
    for (int64_t k = 0; k < n; k++) out[k] = boxmuller(s, irep - 1);
}
void orc_genrand(orc_sys *s, int32_t irep, int32_t kind, int64_t n, double *out) {
    tape_t *t = irep < 0 ? &s->tape_rep : &s->rep[irep - 1].tape;
    for (int64_t k = 0; k < n; k++) out[k] = kind == 1 ? gen_d1(t) : gen_d3(t);
}
void orc_labels(orc_sys *s, int32_t *rep2lab) { for (int k = 1; k <= s->nrep; k++) rep2lab[k - 1] = s->rep2lab[k]; }

/* set_md_coef (src/job_md.F90:700-767): a, sqrt(b)h/m, sqrt(b)h, and the
 * per-replica noise amplitude sqrt(b)/2m*sqrt(2 gamma kT h); for the
 * Brownian integrator (lines 725-731, 755-761) a = h/gamma and the noise
 * amplitude is sqrt(2 kT h / gamma), cf = cx = 0 */
static void md_coef(orc_sys *s, int i, double tK, double *a, double *cf, double *cx, double *cr) {
    if (s->p.integrator == 2) {
        *a = s->p.dt / s->fric[i];
        *cf = 0.0;
        *cx = 0.0;
        *cr = sqrt(2.0 * BOLTZ_KCAL_MOL * tK * s->p.dt / s->fric[i]);
        return;
    }
    double c1 = 0.5 * s->p.dt * s->fric[i] / s->mass[i];
    double c2 = sqrt(1.0 / (1.0 + c1));
    *a = (1.0 - c1) / (1.0 + c1);
    *cf = c2 * s->p.dt / s->mass[i];
    *cx = c2 * s->p.dt;
    *cr = 0.5 * c2 / s->mass[i] * sqrt(2.0 * s->fric[i] * BOLTZ_KCAL_MOL * tK * s->p.dt);
}
void orc_md_coef(orc_sys *s, int32_t irep, double *coef4) {
    for (int i = 0; i < s->nmp; i++)
        md_coef(s, i, s->rep[irep - 1].tK, coef4 + 4 * i, coef4 + 4 * i + 1, coef4 + 4 * i + 2, coef4 + 4 * i + 3);
}

/* initial velocities and accelerations (src/job_md.F90:157-176) */
/* energy_kinetic (src/energy_kinetic.F90:16-29): 0.5 sum_i m_i |v_i|^2,
 * beads in order. */
double orc_kinetic(orc_sys *s, int32_t irep1) {
    const double *v = s->velos + (size_t)(irep1 - 1) * s->nmp * 3;
    double e = 0.0;
    for (int i = 0; i < s->nmp; i++)
        e += s->mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    return 0.5 * e;
}

void orc_init_velocities(orc_sys *s) {
    for (int r = 0; r < s->nrep; r++) {
        double tK = s->rep[r].tK;
        double *v = s->velos + (size_t)r * s->nmp * 3, *ac = s->accels + (size_t)r * s->nmp * 3;
        if (s->ctr_on) {   /* the engine's counter-based draws, same formulas */
            for (int i = 0; i < s->nmp; i++) {
                double nv[3], na[3], a, cf, cx, cr;
                ctr_init_normals(s, r, i, 0, nv);
                ctr_init_normals(s, r, i, 1, na);
                double c1 = sqrt(tK * BOLTZ_KCAL_MOL / s->mass[i]);
                md_coef(s, i, tK, &a, &cf, &cx, &cr);
                for (int k = 0; k < 3; k++) { v[3 * i + k] = c1 * nv[k]; ac[3 * i + k] = cr * na[k]; }
            }
            continue;
        }
        for (int i = 0; i < s->nmp; i++) {
            double c1 = sqrt(tK * BOLTZ_KCAL_MOL / s->mass[i]);
            for (int k = 0; k < 3; k++) v[3 * i + k] = c1 * boxmuller(s, r);
        }
        for (int i = 0; i < s->nmp; i++) {
            double a, cf, cx, cr;
            md_coef(s, i, tK, &a, &cf, &cx, &cr);
            for (int k = 0; k < 3; k++) ac[3 * i + k] = cr * boxmuller(s, r);
        }
    }
}

static int get_pair(orc_sys *s, int lab) { return s->pair_tb[s->idx_pair * (s->nrep + 2) + lab]; }

/* energy_replica + replica_exchange for a 1-D temperature ladder
 * (src/energy_replica.F90:47-156, src/replica_exchange.F90:35-154,
 * src/job_md.F90:536-567) */
static void exchange(orc_sys *s, double *energies_out) {
    int n = s->nrep;
    double *E = calloc(2 * (n + 1), sizeof(double));
    double e[11];
    for (int r = 0; r < n; r++) {
        orc_energy(s, r + 1, s->rep[r].tK, s->p.nstep_bp_mc > 0, 0, e);
        if (energies_out) memcpy(energies_out + 11 * r, e, sizeof e);
        E[2 * (r + 1)] = e[0];
    }
    /* energies at the partner's temperature; BP state saved/restored */
    for (int r = 0; r < n; r++) {
        rep_t *rr = &s->rep[r];
        int64_t nbp = rr->bp.n / 7;
        uint8_t *st = malloc(nbp + 1); double *en = malloc((nbp + 1) * sizeof(double));
        memcpy(st, rr->bp_status, nbp + 1); memcpy(en, rr->ene_bp, (nbp + 1) * sizeof(double));
        int lab = s->rep2lab[r + 1], opp = get_pair(s, lab);
        if (opp == 0 || opp > n) E[2 * (r + 1) + 1] = 0.0;
        else {
            double tk_own = rr->tK, t_opp = s->lab2val[opp];
            int rgl = s->x_on && s->x.rg && s->rg_remd;
            double rg_own = rgl ? s->x_rg0[r] : 0.0;
            set_ele(s, r, t_opp, s->lab2ion[opp]);   /* src/energy_replica.F90:96-112 */
            if (rgl) s->x_rg0[r] = s->lab2rg[opp];
            int fl = s->x_on && s->dcf_remd;
            if (fl) dcf_set_forces(s, r, s->lab2dcf[opp]);
            orc_energy(s, r + 1, t_opp, s->p.nstep_bp_mc > 0, 0, e);
            E[2 * (r + 1) + 1] = e[0];
            set_ele(s, r, tk_own, rr->ionic);
            if (rgl) s->x_rg0[r] = rg_own;
            if (fl) dcf_set_forces(s, r, s->lab2dcf[lab]);
        }
        memcpy(rr->bp_status, st, nbp + 1); memcpy(rr->ene_bp, en, (nbp + 1) * sizeof(double));
        free(st); free(en);
    }
    double *rnd = malloc(sizeof(double) * (n + 1));
    int nr = 0;
    for (int i = 1; i <= n; i += 2) {
        double u = s->ctr_on ? ctr_exchange_uniform(s, nr) : gen_d1(&s->tape_rep);
        rnd[++nr] = u;
    }
    nr = 0;
    for (int li = 1; li <= n; li++) {
        int lj = get_pair(s, li);
        if (lj < li) continue;
        nr++;
        int ri = s->lab2rep[li], rj = s->lab2rep[lj];
        double tm = s->lab2val[li], tn = s->lab2val[lj];
        double pim = E[2 * ri], pin = E[2 * ri + 1], pjn = E[2 * rj], pjm = E[2 * rj + 1];
        double delta = ((pjm - pim) / tm - (pjn - pin) / tn) / BOLTZ_KCAL_MOL;
        int acc = (delta < 0.0) || (rnd[nr] <= exp(-delta));
        if (acc) {
            s->rep2lab[ri] = lj; s->rep2lab[rj] = li;
            s->lab2rep[li] = rj; s->lab2rep[lj] = ri;
            double sf = sqrt(tn / tm);
            double *vi = s->velos + (size_t)(ri - 1) * s->nmp * 3, *vj = s->velos + (size_t)(rj - 1) * s->nmp * 3;
            for (int k = 0; k < 3 * s->nmp; k++) { vi[k] = sf * vi[k]; vj[k] = vj[k] / sf; }
        }
    }
    s->idx_pair = (s->idx_pair + 1) % s->pair_sets;   /* set_forward, src/var_replica.F90:104-117 */
    for (int r = 0; r < n; r++) {
        s->rep[r].tK = s->lab2val[s->rep2lab[r + 1]];
        s->rep[r].ionic = s->lab2ion[s->rep2lab[r + 1]];
        set_ele(s, r, s->rep[r].tK, s->rep[r].ionic);
        if (s->x_on && s->x.rg && s->rg_remd) s->x_rg0[r] = s->lab2rg[s->rep2lab[r + 1]];   /* src/job_md.F90:564-566 */
        if (s->x_on && s->dcf_remd) dcf_set_forces(s, r, s->lab2dcf[s->rep2lab[r + 1]]);   /* :559-562 */
    }
    free(rnd); free(E);
}

/* job_md: initial lists + energies, then the time loop
 * (src/job_md.F90:196-198,276-291,360-460).  Returns the number of list
 * rebuilds; energies_out (nrep*11) gets the last energies. */
int64_t orc_md(orc_sys *s, int64_t nsteps, int32_t nstep_exchange, double *energies_out) {
    int n = s->nmp, nr = s->nrep;
    int64_t rebuilds = 0;
    for (int r = 1; r <= nr; r++) orc_neighbor_list(s, r);
    memset(s->move, 0, sizeof(double) * 3 * n * nr);
    double e[11];
    for (int r = 1; r <= nr; r++) {
        s->flg_bp_energy = 0;
        orc_energy(s, r, s->rep[r - 1].tK, s->p.nstep_bp_mc > 0, 0, e);
        if (energies_out) memcpy(energies_out + 11 * (r - 1), e, sizeof e);
    }
    for (int r = 0; r < nr; r++) {
        rep_t *rr = &s->rep[r];
        int64_t nbp = rr->bp.n / 7;
        for (int64_t b = 0; b < nbp; b++) rr->bp_status_mc[b] = s->p.nstep_bp_mc > 0 ? rr->bp_status[b] : 1;
    }
    double *forces = malloc(sizeof(double) * 3 * n), *rnd = malloc(sizeof(double) * 3 * n);
    int last_exchange = 0;
    for (int64_t st = 0; st < nsteps; st++) {
        s->istep++;
        /* one flag for the whole replica loop (src/job_md.F90:364-366): a
         * rebuild in replica r also forces MC on replicas r+1.. of this step
         * (src/job_md.F90:392-394) */
        int mc = 0;
        if (s->p.nstep_bp_mc > 0 && s->istep % s->p.nstep_bp_mc == 0) mc = 1;
        for (int r = 0; r < nr; r++) {
            double *x = s->xyz + (size_t)r * n * 3, *v = s->velos + (size_t)r * n * 3;
            double *ac = s->accels + (size_t)r * n * 3, *mv = s->move + (size_t)r * n * 3;
            double d2max = 0.0, d2max2 = 0.0;
            for (int i = 0; i < n; i++) {
                double d2 = dot3(mv + 3 * i, mv + 3 * i);
                if (d2 > d2max) { d2max2 = d2max; d2max = d2; }
                else if (d2 > d2max2) d2max2 = d2;
            }
            if (sqrt(d2max) + sqrt(d2max2) > s->p.nl_margin) {
                orc_neighbor_list(s, r + 1);
                memset(mv, 0, sizeof(double) * 3 * n);
                rebuilds++;
                if (s->p.nstep_bp_mc > 0) mc = 1;
            }
            orc_force(s, r + 1, mc, forces);
            if (mc) memcpy(s->rep[r].bp_status_mc, s->rep[r].bp_status, s->rep[r].bp.n / 7 + 1);
            if (s->ctr_on)
                for (int i = 0; i < n; i++) ctr_step_normals(s, r, i, rnd + 3 * i);
            else
                for (int i = 0; i < 3 * n; i++) rnd[i] = boxmuller(s, r);
            for (int i = 0; i < n; i++) {
                double a, cf, cx, cr;
                md_coef(s, i, s->rep[r].tK, &a, &cf, &cx, &cr);
                if (s->p.integrator == 2) {
                    /* Ermak-McCammon (src/job_md.F90:446-456) */
                    for (int k = 0; k < 3; k++) {
                        double dx = a * forces[3 * i + k] + cr * rnd[3 * i + k];
                        x[3 * i + k] = x[3 * i + k] + dx;
                        mv[3 * i + k] = mv[3 * i + k] + dx;
                    }
                    continue;
                }
                for (int k = 0; k < 3; k++) {
                    double apre = cr * rnd[3 * i + k];
                    v[3 * i + k] = a * v[3 * i + k] + cf * forces[3 * i + k] + (ac[3 * i + k] + apre);
                    ac[3 * i + k] = apre;
                    double dx = cx * v[3 * i + k];
                    x[3 * i + k] = x[3 * i + k] + dx;
                    mv[3 * i + k] = mv[3 * i + k] + dx;
                }
            }
        }
        last_exchange = 0;
        if (nr > 1 && nstep_exchange > 0 && s->istep % nstep_exchange == 0) {
            exchange(s, energies_out);
            last_exchange = 1;
        }
        /* energies of the last step as written at a save step, before the
         * resize block: the BP state of the last force call is reused
         * (src/job_md.F90:474-481) */
        if (st == nsteps - 1 && !last_exchange && energies_out) {
            int mc_last = s->p.nstep_bp_mc > 0 && s->istep % s->p.nstep_bp_mc == 0;
            for (int r = 1; r <= nr; r++) orc_energy(s, r, s->rep[r - 1].tK, mc_last, 1, energies_out + 11 * (r - 1));
        }
        /* resize block (src/job_md.F90:576-596) */
        if (s->rz_on && s->istep % s->rz_step == 0) {
            pbc_resize(s);
            for (int r = 0; r < nr; r++) {
                orc_neighbor_list(s, r + 1);
                memset(s->move + (size_t)r * n * 3, 0, sizeof(double) * 3 * n);
            }
        }
        if (s->an_n > 0 && s->an_next > 0 && s->istep + 1 == s->an_next) {
            s->ianneal++;
            for (int r = 0; r < nr; r++) s->rep[r].tK = s->an_T[s->ianneal - 1];
            s->an_next = s->ianneal < s->an_n ? s->an_step[s->ianneal] : 0;
        }
        if (s->rgs_n > 0 && s->rgs_next > 0 && s->istep + 1 == s->rgs_next) {
            s->irgs++;
            rg_schedule_apply(s);
        }
    }
    free(forces); free(rnd);
    return rebuilds;
}
