/*
 * sis_gpu.h - C ABI of the MI355X engine for the SIS molecular-dynamics step.
 *
 * The reference SIS program (Fortran) has no plugin API; its hot path is a
 * set of external subroutines that share module globals.  Each entry point
 * below replaces one of them; the Fortran host binds them with iso_c_binding
 * (sis_amd/fortran/sis_gpu_iface.F90).  All arrays are plain host pointers;
 * the library copies what it needs and never keeps a host pointer.
 *
 * Conventions
 *   - every call returns int: 0 = success, < 0 = error code below;
 *     sis_gpu_last_error() gives the message of the last failure (the
 *     Fortran side maps a failure to sis_abort, src/sis_abort.F90:1-20);
 *   - replica indices irep are 1-based as in the reference (0 = all);
 *   - coordinate arrays use the reference layout xyz(3, nmp, nrep),
 *     Fortran column-major (src/var_state.F90:26-28);
 *   - pair and base-pair lists use the reference layout wca_mp(2, n),
 *     ele_mp(2, n), bp_mp(7, n) with 1-based bead ids
 *     (src/var_potential.F90:86-134);
 *   - one host thread drives a context; calls on one context do not overlap.
 */
#ifndef SIS_GPU_H
#define SIS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SIS_OK = 0,
    SIS_ERR_ARG = -1,         /* invalid argument or unsupported option   */
    SIS_ERR_HIP = -2,         /* HIP runtime failure                      */
    SIS_ERR_UNSUPPORTED = -3, /* feature the GPU backend does not cover   */
    SIS_ERR_CAPACITY = -4,    /* a list outgrew its allocation            */
    SIS_ERR_STATE = -5        /* call order (e.g. force before rebuild)   */
};

/* One base-pair type; field order of basepair_parameters
 * (src/var_potential.F90:63-80).  cutoff_ddist is derived by the library
 * as init_bp does (src/init_bp.F90:321-346). */
typedef struct sis_bp_para {
    double cutoff_ddist, U0, bond_k, bond_r;
    double angl_k1, angl_k2, angl_k3, angl_k4;
    double angl_theta1, angl_theta2, angl_theta3, angl_theta4;
    double dihd_k1, dihd_k2, dihd_phi1, dihd_phi2;
} sis_bp_para;

/* Force field and run parameters: what read_input / read_force_field
 * store in var_potential / var_state (src/read_force_field.F90:74-281,
 * src/read_input.F90:280-790).  bp[] holds GC, AU, GU; CG, UA, UG are the
 * mirrored types (angl_k1<->angl_k2, angl_k3<->angl_k4,
 * src/read_force_field.F90:203-207). */
typedef struct sis_params {
    double bond_k, bond_r0;
    double angl_k, angl_t0;
    double dih_k, dih_w, dih_p0;
    double wca_sigma, wca_eps;
    double ele_cutoff, length_per_charge, ionic_strength;
    double bp_cutoff_energy;
    double bp_dH0, bp_dS0, bp_dHend0, bp_dSend0, bp_coef_dG;
    double nn_dH[21], nn_dS[21];     /* NNT order, src/const_idx.F90:101-124 */
    double nnend_dH[6], nnend_dS[6]; /* NNENDT order */
    sis_bp_para bp[3];
    double tempK;
    double dt, viscosity_pas, nl_margin;
    double box[3];
    int32_t angl_reb;          /* 1: restricted bending (angle_ReB), 0: harmonic */
    int32_t dih_type;          /* 0: none, 1: cosine, 2: exp (Gaussian) */
    int32_t ele_on, ele_cutoff_type, ele_excl_bond, ele_excl_angl, dummy_has_charge;
    int32_t bp_model, bp_min_loop, max_bp_per_nt, nstep_bp_mc, nnend;
    int32_t pbc;
    int32_t integrator;        /* INTGRT (src/const_idx.F90:51-57): 1 GJF-2GJ Langevin,
                                  2 Ermak-McCammon Brownian; 0 means 1 */
} sis_params;

/* Topology: what read_fasta / init_sequence build (src/var_top.F90). */
typedef struct sis_topo {
    int32_t nchains, nmp;
    const int32_t *nmp_chain;  /* [nchains] */
    const int32_t *moltype;    /* [nchains] MOLT: 1 linear, 2 circular */
    const int32_t *seq;        /* [nmp] SEQT codes A0 U1 G2 C3 D4, chain after chain */
    const int32_t *no_charge;  /* [n_no_charge] 1-based beads without charge, or NULL */
    int32_t n_no_charge, pad;
} sis_topo;

/* Engine options (no counterpart in the reference). */
typedef struct sis_options {
    int32_t nrep;            /* replicas (temperature ladder set by set_rep_conditions) */
    int32_t precision;       /* 0: FP64 everywhere; 1: FP32 pair arithmetic, FP64 state */
    int32_t parity;          /* 1: consume injected MT19937 words as the reference does */
    int32_t mode;            /* 0 auto, 1 one launch per phase, 2 persistent, 3 batched */
    int32_t block;           /* batched form: threads per replica's workgroup, 128, 256 or 512
                                (0 = 256); other values are refused */
    int32_t max_nbr;         /* neighbour slots per bead, 0 = grow on demand */
    int32_t rebuild_every;   /* 0: displacement trigger (src/job_md.F90:376-395) */
    int32_t exchange_every;  /* replica exchange cadence in steps, 0 = none */
    uint64_t seed;           /* counter-RNG seed for the performance mode */
    int32_t grid;            /* persistent form: workgroups of the cooperative launch
                                (0 = as many as fit on the device) */
    int32_t flags;           /* bit0 SIS_OPT_HOST_EXCHANGE: batched runs exchange on the
                                host between segments (no device exchange, no graph);
                                bit1 SIS_OPT_NO_GRAPH: the per-phase form launches every
                                step from the host instead of replaying a captured graph */
} sis_options;
#define SIS_OPT_HOST_EXCHANGE 1
#define SIS_OPT_NO_GRAPH 2

typedef struct sis_gpu_ctx sis_gpu_ctx;

/* Set up a system; derives what init_bp, init_ele, list_local and
 * set_md_coef derive (src/init_bp.F90, src/set_ele.F90:45-99,
 * src/list_local.F90, src/job_md.F90:94-142,700-767). */
int sis_gpu_create(const sis_topo *topo, const sis_params *prm,
                   const sis_options *opt, sis_gpu_ctx **ctx);
int sis_gpu_destroy(sis_gpu_ctx *ctx);
const char *sis_gpu_last_error(void);

/* Copy bead state in / out; xyz, velos, accels are (3, nmp, nrep);
 * velos / accels may be NULL. */
int sis_gpu_set_state(sis_gpu_ctx *ctx, const double *xyz, const double *velos,
                      const double *accels);
int sis_gpu_get_state(sis_gpu_ctx *ctx, double *xyz, double *velos, double *accels);

/* Coordinates (3, nmp) of replica irep, as the reference's force /
 * energy_sumup read them from var_state's xyz(:,:,irep) when its own loop
 * integrates on the host (src/job_md.F90:397,290).  Unlike
 * sis_gpu_set_state the lists are kept: the caller rebuilds them when
 * its own skin test says so (src/job_md.F90:372-395). */
int sis_gpu_set_positions(sis_gpu_ctx *ctx, int32_t irep, const double *xyz);

/* Per-replica temperature [nrep] and ionic strength [nrep] (NULL keeps the
 * current value); re-derives set_ele and set_md_coef per replica
 * (src/job_md.F90:536-567). */
int sis_gpu_set_rep_conditions(sis_gpu_ctx *ctx, const double *tempK, const double *ionic);

/* neighbor_list(irep): pbc_wrap + WCA / DH / BP lists
 * (src/neighbor_list.F90:1-233, called from src/job_md.F90:197,389,584). */
int sis_gpu_rebuild(sis_gpu_ctx *ctx, int32_t irep);

/* The lists of the jobs other than MD (src/main.F90:144-154: list_bp,
 * list_exv, list_ele; src/list_bp.F90:41-93, src/list_exv.F90:19-64,
 * src/list_ele.F90:18-65) for every replica: no circular wrap in the
 * exclusions, list_ele's bond / angle exclusions on every chain pair,
 * base-pair candidates with jmp >= imp + bp_min_loop + 1 over the whole
 * system in (imp, jmp) order.  Force and energy calls then use them; the
 * next step rebuilds the neighbour lists of the MD job.  Pairs beyond the
 * neighbour cutoffs are not listed (they have no energy, force or formed
 * base pair). */
int sis_gpu_analysis_lists(sis_gpu_ctx *ctx);

/* force(irep, forces) (src/force.F90:1-77); bp_mc = flg_bp_MC.
 * forces is (3, nmp) or NULL (keep on device). */
int sis_gpu_force(sis_gpu_ctx *ctx, int32_t irep, int32_t bp_mc, double *forces);

/* energy_sumup(irep, tempK, energies) (src/energy_sumup.F90:1-56);
 * flags bit0 = flg_bp_MC, bit1 = flg_bp_energy (reuse the BP state of the
 * last force call).  energies[11] in ENE order (src/const_idx.F90:21). */
int sis_gpu_energy(sis_gpu_ctx *ctx, int32_t irep, double tempK, int32_t flags,
                   double *energies);

/* sis_gpu_energy without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and summation order as sis_gpu_energy; its partial sums, output
 * row and per-call conditions live in a static device array of
 * SIS_NOALLOC_SCRATCH_BYTES per device instead of library allocations.
 * Needs (8*ceil(nmp/256) + 16 + nrep) doubles + 2*nrep ints of it;
 * larger contexts get SIS_ERR_CAPACITY.  flags: bit1 as in
 * sis_gpu_energy; bit0 (the limit MC) is refused.  The context's device
 * must be current.  Every call on a device shares that device's array:
 * calls that overlap on one device race. */
#define SIS_NOALLOC_SCRATCH_BYTES 262144
int sis_gpu_energy_noalloc(sis_gpu_ctx *ctx, int32_t irep, double tempK, int32_t flags,
                           double *energies);

/* Parity dumps of the interaction lists.  kind 0 = wca_mp, 1 = ele_mp
 * (2 ints per pair, i < j), 2 = bp_mp (7 ints per candidate, list order
 * of src/neighbor_list.F90:99-114). */
int sis_gpu_get_pair_count(sis_gpu_ctx *ctx, int32_t irep, int32_t kind, int64_t *n);
int sis_gpu_get_pairs(sis_gpu_ctx *ctx, int32_t irep, int32_t kind, int32_t *out,
                      int64_t cap, int64_t *n);
/* BP candidates with bp_coef(2, n), bp_status and ene_bp of the last force
 * or energy call (src/force_bp_limit_triplet.F90:243-250). */
int sis_gpu_get_bp(sis_gpu_ctx *ctx, int32_t irep, int32_t *bp_mp, double *bp_coef,
                   uint8_t *status, double *ene, int64_t cap, int64_t *n);

/* Random numbers.  Performance mode: counter-based normals from `seed`.
 * Parity mode: replica irep consumes `words` (tempered MT19937 output of
 * its mts(irep) stream) exactly as rnd_boxmuller / genrand_double1 /
 * genrand_double3 would (src/rnd_boxmuller.F90:1-49,
 * mt_stream.F90:708-775); irep = -1 gives the exchange stream mts_rep. */
int sis_gpu_set_rng(sis_gpu_ctx *ctx, uint64_t seed);
int sis_gpu_set_rng_tape(sis_gpu_ctx *ctx, int32_t irep, const uint32_t *words, int64_t n);

/* Parity mode: words of replica irep's tape used so far (irep = -1: the
 * exchange tape), so that a caller can advance its own mts(irep) stream
 * by what the limit MC drew (genrand_double1 / genrand_double3 in
 * src/force_bp_limit_triplet.F90, src/energy_bp_limit_triplet.F90). */
int sis_gpu_get_rng_tape_pos(sis_gpu_ctx *ctx, int32_t irep, int64_t *pos);

/* Maxwell-Boltzmann velocities and initial GJF noise terms
 * (src/job_md.F90:157-176). */
int sis_gpu_init_velocities(sis_gpu_ctx *ctx);

/* nsteps of the job_md loop body for every replica: skin test and
 * rebuild, force (with the BP limit MC every nstep_bp_mc steps), noise,
 * GJF-2GJ or Ermak-McCammon update (src/job_md.F90:360-460), and replica exchange every
 * exchange_every steps (src/job_md.F90:462-570).  *rebuilds (may be NULL)
 * returns the number of neighbour-list rebuilds done. */
int sis_gpu_step(sis_gpu_ctx *ctx, int64_t nsteps, int64_t *rebuilds);

/* energy_replica + replica_exchange now (src/energy_replica.F90,
 * src/replica_exchange.F90); replica_energies(2, nrep) out (may be NULL),
 * accepted[nrep/2] out (may be NULL). */
int sis_gpu_exchange(sis_gpu_ctx *ctx, double *replica_energies, int32_t *accepted);

/* Replicas on a grid of ndim dimensions with nvalues[d] values along each
 * (src/init_replica.F90:39-93: the variables with ladders in REPT order,
 * temperature, force, ionic strength, Rg; nrep = the product; label l's
 * values laid out with the first dimension slowest, as lab2val).  Builds
 * exchange_pair_tb (src/var_replica.F90:162-259): exchanges take the odd
 * pairs along each dimension in turn, then the even ones, and start again
 * (set_forward, :104-117).  Without this call the replicas are one ladder
 * (ndim = 1).  The labels' conditions are set with
 * sis_gpu_set_rep_conditions / sis_gpu_set_rep_rg / sis_gpu_set_rep_dcf. */
int sis_gpu_set_rep_dims(sis_gpu_ctx *ctx, int32_t ndim, const int32_t *nvalues);
/* rep2lab(1:nrep) (src/var_replica.F90:49). */
int sis_gpu_get_labels(sis_gpu_ctx *ctx, int32_t *rep2lab);

/* External terms beside the bonded and pair potentials
 * (src/force.F90:62-75, src/energy_sumup.F90:46-53): the wall at z = 0
 * (force_stage / energy_stage, src/force_stage.F90, src/energy_stage.F90),
 * optical tweezers with constant-force pairs and force ramps
 * (src/force_tweezers.F90, src/energy_tweezers.F90; the ramp anchors move
 * as init + velo * istep), the Rg bias (src/force_rg.F90,
 * src/energy_rg.F90) and sigmoid restraints (src/force_restraint.F90,
 * src/energy_restraint.F90).  Arrays are in the reference's layout with
 * 1-based bead ids; energies go to ENE STAGE, TWZ, RG and REST (7-10). */
typedef struct sis_external {
    int32_t stage;                /* flg_stage */
    int32_t n_dcf;                /* ntwz_DCF */
    int32_t n_fr;                 /* ntwz_FR */
    int32_t rg;                   /* 0 off, else bias_rg_pott (POTT: 1 harmonic, 2 flat-bottomed) */
    int32_t n_rest;               /* nrest_sigb */
    int32_t pad;
    double stage_sigma, stage_eps;
    double rg_k;                  /* bias_rg_k */
    const double *rg_0;           /* bias_rg_0 [nrep] */
    const int32_t *dcf_pairs;     /* twz_DCF_pairs (2, n_dcf) */
    const double *dcf_forces;     /* twz_DCF_forces (3, n_dcf, nrep) */
    const int32_t *fr_pairs;      /* twz_FR_pairs (2, n_fr) */
    const double *fr_k;           /* twz_FR_k (2, n_fr) */
    const double *fr_init;        /* twz_FR_init (3, 2, n_fr) */
    const double *fr_velo;        /* twz_FR_velo (3, 2, n_fr) */
    const int32_t *rest_id;       /* rest_sigb_id (3, n_rest): bead, reference bead, 1 sigmoid / 2 sigmoid-to-bead */
    const double *rest_rcut;      /* rest_sigb_rcut (n_rest) */
    const double *rest_para;      /* rest_sigb_para (3, n_rest): epsilon, r_bound, delta */
} sis_external;

/* Set (or with NULL / all terms off, clear) the external terms.  They run
 * in the per-phase form; a context with external terms steps in that form
 * whatever mode it was created with. */
int sis_gpu_set_external(sis_gpu_ctx *ctx, const sis_external *ext);

/* Box resizing ([PBC_box] resize_step / resize_change / resize_target;
 * pbc_resize, src/pbc.F90:79-138, applied by src/job_md.F90:576-596):
 * every resize_step steps each axis becomes original + change * count
 * (original = the box when this is called), clamped at target (NULL: no
 * target) until every axis reaches it; then every replica's lists are
 * rebuilt and later DCD frames and restart files carry the new box.
 * resize_step <= 0 turns it off.  Needs a periodic box. */
int sis_gpu_set_box_resize(sis_gpu_ctx *ctx, int64_t resize_step, const double *change, const double *target);

/* Replica exchange over the Rg-bias target (REPT%RG; src/init_bias_rg.F90:13-16,
 * src/energy_replica.F90:107-109, src/job_md.F90:564-566): rg0[nrep] are
 * the targets of the replicas' current labels and travel with them
 * through exchanges.  Needs the Rg bias of sis_gpu_set_external (which
 * drops the ladder again); NULL stops it. */
int sis_gpu_set_rep_rg(sis_gpu_ctx *ctx, const double *rg0);

/* Replica exchange over the tweezers' constant force (REPT%TWZDCF,
 * f-REMD; src/init_tweezers.F90:24-40, src/energy_replica.F90:103-105,
 * src/job_md.F90:559-562): one constant-force pair whose force is
 * k[label] times the normalised direction[3]; k travels with the labels.
 * NULL k stops it. */
int sis_gpu_set_rep_dcf(sis_gpu_ctx *ctx, const double *k, const double *direction);

/* Timed Rg bias (read_timed_bias_rg, src/read_timed_bias_rg.F90;
 * src/job_md.F90:231-265,614-627): entry i sets the Rg-bias constant k[i]
 * and target r0[i] of every replica from step steps[i] on.  Needs the Rg
 * bias of sis_gpu_set_external; a later sis_gpu_set_external drops the
 * schedule, so call this after it.  n <= 0 stops it. */
int sis_gpu_set_rg_schedule(sis_gpu_ctx *ctx, int32_t n, const int64_t *steps, const double *k, const double *r0);

/* Simulated annealing (read_anneal's schedule, src/read_anneal.F90;
 * src/main.F90:90-97, src/job_md.F90:207-230,600-612): entry k sets the
 * temperature temps[k] from step steps[k] on (steps increasing; the first
 * entry is the starting temperature).  As in the reference, the noise
 * amplitude, the base-pair dG and the energies follow the schedule, the
 * electrostatics stay at the first entry's temperature and the force's
 * limit MC keeps the kT of sis_params.tempK.  Call it after
 * sis_gpu_set_state / sis_gpu_rst_read and before sis_gpu_init_velocities;
 * n <= 0 turns it off.  Not with temperature exchange (nrep > 1 and
 * exchange_every > 0). */
int sis_gpu_set_anneal(sis_gpu_ctx *ctx, int32_t n, const int64_t *steps, const double *temps);

/* Secondary-structure restriction (read_ss, src/read_ss.F90, applied by
 * init_bp, src/init_bp.F90:177-209): only the listed pairs (npair pairs of
 * 1-based bead ids, either order, as a CT, BPSEQ or BPL file gives them)
 * may form, after read_ss's filters (chain ends, minimum loop, isolated
 * pairs for models 3-5, non-canonical pairs).  npair < 0 lifts the
 * restriction.  The lists are rebuilt by the next step.  The distance bias
 * of flg_bias_ss (src/force_bp_limit_triplet.F90:104-110) is not covered. */
int sis_gpu_set_ss(sis_gpu_ctx *ctx, int32_t npair, const int32_t *pairs);

/* Restart: the STEP and REPLICA blocks of a restart file
 * (src/write_rst.F90:40-60, read back by src/read_rst.F90).  set_step
 * sets istep, which keys the performance-mode random streams, so a run
 * restarted from sis_gpu_get_state / get_labels / get_info continues the
 * same streams.  set_labels takes rep2lab(1:nrep) after
 * sis_gpu_set_rep_conditions gave the temperature of each label; each
 * replica takes its label's temperature, and the exchange pair table
 * starts from its first half as the reference's does after a restart. */
int sis_gpu_set_step(sis_gpu_ctx *ctx, int64_t istep);
int sis_gpu_set_labels(sis_gpu_ctx *ctx, const int32_t *rep2lab);

/* energy_kinetic (src/energy_kinetic.F90:1-31): 0.5 sum m |v|^2 of
 * replica irep (1-based), or of every replica into ekin[nrep] when
 * irep = 0. */
int sis_gpu_kinetic(sis_gpu_ctx *ctx, int32_t irep, double *ekin);

/* Trajectory output in the reference's DCD layout (file_dcd of
 * src/dcd.F90: write_header :287-355, write_onestep :538-596), one file
 * per replica as job_md opens them (src/job_md.F90:294-303).  The unit
 * cell blocks are written when the system is periodic; fix_com_origin > 0
 * subtracts the centre of geometry as write_onestep does.  With
 * nstep_save > 0, sis_gpu_step writes a frame of every replica after each
 * step whose count is a multiple of nstep_save (the save block of
 * src/job_md.F90:485-508): the positions are copied to pinned host memory
 * on the engine's stream and written by a background thread while the
 * next steps run.  sis_gpu_dcd_write writes the current state now (the
 * initial frame of src/job_md.F90:336-339) and returns when it is on disk;
 * sis_gpu_dcd_close waits for queued frames and closes the files (also
 * done by sis_gpu_destroy).  paths holds nrep file names. */
int sis_gpu_dcd_open(sis_gpu_ctx *ctx, const char *const *paths, int32_t fix_com_origin, int64_t nstep_save);
int sis_gpu_dcd_write(sis_gpu_ctx *ctx);
int sis_gpu_dcd_close(sis_gpu_ctx *ctx);

/* Base-pair output in the layout of src/write_bp.F90 (header of
 * src/init_out_files.F90:102-104): kind 0 = .bp (formed pairs with
 * E <= -kT, src/write_bp.F90:28-51), kind 1 = .bpall (all formed pairs,
 * :53-82); records (int16 i, int16 j, float32 E) in list order and a
 * (0, 0, 0.0) terminator per save step.  With nstep_save > 0 sis_gpu_step
 * writes every replica's records at the save steps, before that step's
 * exchange (the save block precedes the exchange block in
 * src/job_md.F90:485-570); sis_gpu_bp_write writes the current state (the
 * initial write_bp of src/job_md.F90:336), which is that of the last
 * force or energy evaluation. */
int sis_gpu_bp_open(sis_gpu_ctx *ctx, const char *const *paths, int32_t kind, int64_t nstep_save);
int sis_gpu_bp_write(sis_gpu_ctx *ctx);
int sis_gpu_bp_close(sis_gpu_ctx *ctx);

/* Energy lines (.out) in the format of src/write_out.F90: the header of
 * write_out_header and, per save step, istep, the replica label (REMD),
 * T, Ekin and energies(0:5) [and Eele] with the edit descriptors i12, i4,
 * f6.2 and g13.6.  With nstep_save > 0 sis_gpu_step writes a line per
 * replica at the save steps: labels, T and Ekin before that step's
 * exchange, energies at the replica's own temperature (with a fresh limit
 * MC on exchange steps, reusing the step's base-pair state otherwise, as
 * src/job_md.F90:474-510 does). */
int sis_gpu_out_open(sis_gpu_ctx *ctx, const char *const *paths, int64_t nstep_save);
int sis_gpu_out_close(sis_gpu_ctx *ctx);
/* The lines of the current state now: the initial write_out of
 * src/job_md.F90:280-291,331-333 (energies with a fresh limit MC). */
int sis_gpu_out_write(sis_gpu_ctx *ctx);

/* Tweezers' force-ramp lines (.twz, src/init_out_files.F90:133-139;
 * written by force_tweezers, src/force_tweezers.F90:50-59): one file, its
 * header, then at every step whose count is a multiple of nstep_save one
 * line per replica and force-ramp pair (step, pair, trap distance, bead
 * distance, bead-trap distances, forces in pN) from the positions at the
 * start of that step. */
int sis_gpu_twz_open(sis_gpu_ctx *ctx, const char *path, int64_t nstep_save);
int sis_gpu_twz_close(sis_gpu_ctx *ctx);

/* Restart files (.rst) in the block format of write_rst
 * (src/write_rst.F90:25-145; block ids src/const_idx.F90:72-85), one per
 * replica: REPLICA (replica 1's file, nrep > 1), STEP, XYZ, VELO, ACCEL,
 * PBC (replica 1's file, periodic systems) and block 11, the counter-RNG
 * seed of the performance mode (uint64; readers that do not ask for it
 * skip it, src/read_rst.F90:48-53).  The parity form's MT streams belong
 * to the caller, so the device-fed files hold no PRNG / PRNGREP blocks
 * (read_rst then reports them absent, src/read_rst.F90:640-660).  With
 * nstep_save_rst > 0, sis_gpu_step rewrites the files after each step whose
 * count is a multiple of it, after that step's exchange
 * (src/job_md.F90:664-666); sis_gpu_rst_write writes them now.
 * sis_gpu_rst_read is read_rst for a restarted run (src/read_rst.F90, called
 * from src/job_md.F90:149-150,202, src/init_replica.F90:148, src/pbc.F90:35-45):
 * it restores labels, state, step count and the PBC block's box; the lists
 * are rebuilt by the next step. */
int sis_gpu_rst_open(sis_gpu_ctx *ctx, const char *const *paths, int64_t nstep_save_rst);
int sis_gpu_rst_write(sis_gpu_ctx *ctx);
int sis_gpu_rst_close(sis_gpu_ctx *ctx);
int sis_gpu_rst_read(sis_gpu_ctx *ctx, const char *const *paths);

/* job_dcd (src/job_dcd.F90:59-80): each frame of the DCD file at path
 * (read as src/dcd.F90 read_onestep does, float32 to double) becomes the
 * coordinates of replica 1; the lists of sis_gpu_analysis_lists are built
 * and the energies
 * evaluated with the limit MC when nstep_bp_MC > 0.  energies (11,
 * max_frames) receives them, *nframes the frame count; with out_path the
 * .out lines of job_dcd are written (frame number, T, Ekin = 0). */
int sis_gpu_dcd_energies(sis_gpu_ctx *ctx, const char *path, const char *out_path, int64_t max_frames,
                         double *energies, int64_t *nframes);

/* job_check_force (src/job_check_force.F90:46-82): forces of replica irep
 * (f_force, 3 x nmp) against central differences of its total energy with
 * step h (f_energy, 3 x nmp), on the lists of sis_gpu_analysis_lists; the
 * limit MC as the run is configured.  With
 * out_path the per-bead lines of job_check_force are written. */
int sis_gpu_check_force(sis_gpu_ctx *ctx, int32_t irep, double h, const char *out_path, double *f_energy,
                        double *f_force);

/* DCD reader on the host: up to max_frames frames into xyz (3, nmp,
 * max_frames); *nframes = frames in the file; box (3) = the last unit
 * cell, or -1 when the file has none. */
int sis_dcd_read_file(const char *path, int32_t nmp, int64_t max_frames, double *xyz, double *box,
                      int64_t *nframes);

/* The restart writer on host arrays for replica grep: xyz, velos, accels
 * (3, nmp); rep2lab [nrep] (replica 1 with nrep > 1); mts_rep / mts the
 * 643-word mt_stream states of the PRNGREP / PRNG blocks, box[3] (PBC
 * block) and seed (block 11) may each be NULL to leave the block out.
 * Needs no GPU. */
int sis_rst_write_file(const char *path, int32_t grep, int32_t nrep, int32_t nmp, int64_t istep,
                       const int32_t *rep2lab, const double *xyz, const double *velos, const double *accels,
                       const uint32_t *mts_rep, const uint32_t *mts, const double *box, const uint64_t *seed);

/* Parse a restart file: *blocks gets bit (id-1) for every block present;
 * each output may be NULL.  Needs no GPU. */
int sis_rst_read_file(const char *path, int32_t nmp, int32_t nrep, int32_t *blocks, int32_t *grep,
                      int64_t *istep, int32_t *rep2lab, double *xyz, double *velos, double *accels,
                      double *box, uint64_t *seed);

/* The same line writer on host arrays: energies (11, nline). */
int sis_out_write_file(const char *path, int32_t nline, const int64_t *isteps, const int32_t *labels,
                       const double *temps, const double *ekin, const double *energies, int32_t replica,
                       int32_t ele);

/* The same writer on host arrays: nsets save steps of n candidates each,
 * ij (2, n) 1-based bead ids, status (n) and ene (n) per set. */
int sis_bp_write_file(const char *path, int64_t n, const int32_t *ij, const uint8_t *status, const double *ene,
                      int32_t nsets, double tempK, int32_t kind);

/* The same DCD writer on host arrays: nframes frames of xyz(3, nmp)
 * (Fortran layout, frames consecutive); box[3] is used when unitcell != 0.
 * Needs no GPU. */
int sis_dcd_write_file(const char *path, int32_t nmp, int32_t nframes, const double *xyz, int32_t unitcell,
                       const double *box, int32_t fix_com_origin);

/* Project-written autotuner (no reference counterpart).  From the current
 * state, runs nsteps steps of the context's form for every launch shape
 * (batched form: 128, 256 and 512 threads per replica; persistent form:
 * 1 and 2 workgroups per compute unit and as many as fit; per-phase form
 * with device-driven batches of steps: 0 replayed from a captured HIP
 * graph, 1 the same launches issued directly) and every
 * neighbour-list skin skins[0..nskin) (A), restoring the state before each
 * trial, and keeps the fastest shape and skin (the skin moves rebuilds,
 * not forces).  The state is restored at the end and the lists are
 * rebuilt at the next step.  report (may be NULL): [0] chosen shape,
 * [1] chosen skin, then per trial (shape, skin, ms per step);
 * 2 + 9 * nskin doubles. */
int sis_gpu_autotune(sis_gpu_ctx *ctx, const double *skins, int32_t nskin, int64_t nsteps, double *report);

/* Counters: out[0] nmp, [1] nrep, [2] istep, [3] rebuilds, [4] nwca,
 * [5] nele, [6] nbp (replica 1), [7] neighbour slots per bead,
 * [8] launch mode used, [9] workgroup size (batched form: threads per
 * replica; persistent form: workgroups), [10] exchanges done on the
 * device. */
int sis_gpu_get_info(sis_gpu_ctx *ctx, int64_t *out, int32_t n);
/* Device time (ms, HIP events on the context stream) of the last call of
 * a kind: 0 = step, 1 = rebuild, 2 = force; 3 = mean duration of one
 * launch of the step's dominant kernel during the last sis_gpu_step call
 * (the fused force + integrator kernel, or the persistent / batched step
 * kernel), 4 = number of such launches. */
double sis_gpu_last_kernel_ms(sis_gpu_ctx *ctx, int32_t which);

/* ---- Job pool: the serving form -------------------------------------
 * No counterpart in the reference, where one job is one process.  Jobs
 * (own topology, temperature, salt, box, steps and seed; the force field
 * of the pool's first job) are submitted as they arrive.  A scheduler
 * thread admits them in arrival order while their HBM footprint fits
 * hbm_budget_bytes (0 = 80% of free memory).  It keeps their state
 * resident and advances all resident jobs in shared launches, each job by
 * about quantum_bead_steps per round (0 = 2e6).  A job runs the job_md loop
 * body (src/job_md.F90:360-460) in performance mode; its result depends
 * only on its own inputs, not on the other jobs in the pool. */
enum { SIS_JOB_QUEUED = 0, SIS_JOB_RUNNING = 1, SIS_JOB_DONE = 2, SIS_JOB_FAILED = 3 };
typedef struct sis_pool sis_pool;
int sis_pool_create(uint64_t hbm_budget_bytes, int64_t quantum_bead_steps, int32_t max_active,
                    sis_pool **pool);
/* xyz (3, nmp) of one replica; prm->tempK, ionic_strength and box are the job's
 * own.  precision as in sis_options. */
int sis_pool_submit(sis_pool *pool, const sis_topo *topo, const sis_params *prm, const double *xyz,
                    int32_t precision, int64_t nsteps, uint64_t seed, int64_t *job_id);
/* state = SIS_JOB_*; times[5] (may be NULL) = submit, start, end (s since the pool
 * was created, -1 if not yet), steps done, rebuilds; xyz (3, nmp) is filled once
 * the job is done.  A failed job returns SIS_ERR_STATE with its message. */
int sis_pool_poll(sis_pool *pool, int64_t job_id, int32_t *state, double *times, double *xyz);
/* drop the record of a finished or failed job */
int sis_pool_forget(sis_pool *pool, int64_t job_id);
/* out[0..8]: queued, running, done, failed, bead-steps, GPU seconds in launches,
 * launches, HBM bytes held by resident jobs, seconds since creation */
int sis_pool_stats(sis_pool *pool, double *out, int32_t n);
int sis_pool_destroy(sis_pool *pool);

#ifdef __cplusplus
}
#endif

#endif /* SIS_GPU_H */
