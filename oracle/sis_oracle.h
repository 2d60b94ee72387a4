/* CPU restatement of the SIS hot path, used only as a test oracle (tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg).  Not part of
 * the product.  Each function cites the reference routine it restates. */
#ifndef SIS_ORACLE_H
#define SIS_ORACLE_H
#include <stdint.h>
#include "../include/sis_gpu.h"

typedef struct orc_sys orc_sys;

orc_sys *orc_create(const sis_topo *topo, const sis_params *prm, int32_t nrep,
                    const double *temps);
void orc_destroy(orc_sys *s);
void orc_set_xyz(orc_sys *s, const double *xyz);           /* (nrep, nmp, 3) */
void orc_get_state(orc_sys *s, double *xyz, double *velos, double *accels);
double orc_kinetic(orc_sys *s, int32_t irep1);
void orc_set_state(orc_sys *s, const double *xyz, const double *velos, const double *accels);
void orc_neighbor_list(orc_sys *s, int32_t irep);            /* 1-based irep */
int64_t orc_count(orc_sys *s, int32_t irep, int32_t kind);   /* 0 wca 1 ele 2 bp */
void orc_lists(orc_sys *s, int32_t irep, int32_t kind, int32_t *out, double *bpcoef);
void orc_force(orc_sys *s, int32_t irep, int32_t bp_mc, double *forces);
void orc_energy(orc_sys *s, int32_t irep, double tK, int32_t bp_mc, int32_t bp_energy_valid,
                double *energies);
void orc_bp_state(orc_sys *s, int32_t irep, uint8_t *status, double *ene);
void orc_tables(orc_sys *s, double *dh, double *ds, double *cut);
void orc_bp_pair(orc_sys *s, int32_t imp, int32_t jmp, int32_t *bptype, int32_t *bp3);
void orc_ele_params(orc_sys *s, int32_t irep, double *out);
void orc_md_coef(orc_sys *s, int32_t irep, double *coef4);  /* (N,4) a, c_f, c_x, c_r */
/* RNG: tape of tempered MT words per replica (irep = -1: exchange stream) */
void orc_set_tape(orc_sys *s, int32_t irep, const uint32_t *words, int64_t n);
int64_t orc_tape_pos(orc_sys *s, int32_t irep);
void orc_boxmuller(orc_sys *s, int32_t irep, int64_t n, double *out);
void orc_mt_words(uint32_t seed, int64_t n, uint32_t *out);  /* init_genrand stream */
void orc_genrand(orc_sys *s, int32_t irep, int32_t kind, int64_t n, double *out);
/* MD */
void orc_init_velocities(orc_sys *s);
int64_t orc_md(orc_sys *s, int64_t nsteps, int32_t nstep_exchange, double *energies_out);
void orc_labels(orc_sys *s, int32_t *rep2lab);
void orc_set_temps(orc_sys *s, const double *temps);
void orc_set_ionic(orc_sys *s, const double *ionic);
void orc_set_rep_rg(orc_sys *s, const double *rg0);
void orc_set_rep_dcf(orc_sys *s, const double *k, const double *dir);   /* f-REMD */    /* per label: REMD over the Rg target */   /* per label: REMD over the ionic strength */
int64_t orc_mc_iterations(void);
void orc_set_external(orc_sys *s, const sis_external *ext);   /* NULL clears */
void orc_set_step(orc_sys *s, int64_t istep);
void orc_get_box(orc_sys *s, double *box);
void orc_set_box_resize(orc_sys *s, int64_t resize_step, const double *change, const double *target);
void orc_set_rg_schedule(orc_sys *s, int32_t n, const int64_t *steps, const double *k, const double *r0);
void orc_set_anneal(orc_sys *s, int32_t n, const int64_t *steps, const double *temps);
int32_t orc_set_ss(orc_sys *s, int32_t npair, const int32_t *pairs);   /* (2, npair) 1-based; npair < 0 clears */
void orc_plant(orc_sys *s, int32_t which, int32_t on);   /* test hook: planted mistakes */
/* counter-based random numbers of the engine's performance form in place of
 * the MT tapes (Philox4x32-10 keyed by seed, replica, step, bead) */
void orc_set_counter_rng(orc_sys *s, uint64_t seed, int32_t on);
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_set_nl_brute(orc_sys *s, int32_t on);   /* 1: neighbour lists by the O(N^2) loops only */
void orc_analysis_lists(orc_sys *s, int32_t irep);   /* list_bp / list_exv / list_ele (DCD, CHECK_FORCE jobs) */
void orc_set_rep_dims(orc_sys *s, int32_t ndim, const int32_t *nvalues);   /* exchange_pair_tb of a replica grid */
void orc_pair_table(orc_sys *s, int32_t *out);   /* [2 ndim][nrep + 2] */
#endif
