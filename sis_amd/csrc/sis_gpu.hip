// sis_gpu.hip - host side of the MI355X engine: context set-up, list
// management, the per-phase step loop and the C ABI of include/sis_gpu.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sis_gpu.h"
#include "sis_device.h"
#include "sis_kernels.h"
#include "sis_persistent.h"
#include "sis_exchange.h"
#include "sis_external.h"

using namespace sis;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_));                    \
  } while (0)

struct HipError {
  std::string msg;
  explicit HipError(std::string m) : msg(std::move(m)) {}
};
struct ApiError {
  int code;
  std::string msg;
};

const double KB_KCAL = 1.380649e-23 * ((1.0 / 4184.0) * 6.02214076e23);  // src/const_phys.F90:10-14
const double PI_D = 3.141592653589793115997963468544185161590576171875;

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (p) {
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipFree(p));
    }
    p = nullptr;
    n = count;
    if (count) {
      // the engine works on a non-blocking stream: finish the zero fill
      // before any kernel can touch the new buffer
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipMalloc(&p, count * sizeof(T)));
      HIPCHK(hipMemset(p, 0, count * sizeof(T)));
      HIPCHK(hipDeviceSynchronize());
    }
  }
  void free_() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct DcdOut;
struct BpOut;
struct OutFile;
struct RstOut;
struct TwzOut;

// What captured per-phase steps take by value: their graph is replayed only
// while all of it is unchanged.
struct PhaseGraphKey {
  Params P;
  Ptrs D;
  const void *pos;
  int nblk, block;
};

struct sis_gpu_ctx {
  DcdOut *dcd = nullptr;   // trajectory output (sis_gpu_dcd_open), or none
  BpOut *bpo = nullptr;    // base-pair output (sis_gpu_bp_open), or none
  OutFile *outf = nullptr; // energy lines (sis_gpu_out_open), or none
  RstOut *rst = nullptr;   // restart files (sis_gpu_rst_open), or none
  TwzOut *twz = nullptr;   // tweezers' force-ramp lines (sis_gpu_twz_open), or none
  // annealing schedule (sis_gpu_set_anneal): temperatures from step
  // an_step[k]; ele_T holds the temperature the electrostatics were set at
  bool anneal_on = false;
  std::vector<int64_t> an_step;
  std::vector<double> an_T, ele_T;
  int ianneal = 0;            // 1-based index of the current schedule entry
  int64_t an_next = 0;        // step whose start takes the next entry, 0 = none
  int rst_ianneal = 0;        // ANNEAL block of a restart file read, or 0
  // timed Rg-bias schedule (sis_gpu_set_rg_schedule)
  std::vector<int64_t> rgs_step;
  std::vector<double> rgs_k, rgs_0;
  int irgs = 0;
  int64_t rgs_next = 0;       // step whose start takes the next entry, 0 = none
  // box resizing (sis_gpu_set_box_resize; src/pbc.F90:79-138)
  bool rz_on = false, rz_has_target = false;
  int64_t rz_step = 0;
  int64_t rz_count[3] = {0, 0, 0};
  double rz_change[3] = {0, 0, 0}, rz_target[3] = {0, 0, 0}, rz_orig[3] = {0, 0, 0};
  // secondary-structure restriction (sis_gpu_set_ss)
  bool ss_on = false;
  DevBuf<int> ss_off_d, ss_partner_d;
  // external terms (sis_gpu_set_external)
  bool ext_on = false;
  ExtDev X{};
  DevBuf<double> x_rg0_d, x_rg0_opp_d, x_dcf_opp_d, x_dcf_f_d, x_fr_k_d, x_fr_init_d, x_fr_velo_d, x_rest_rcut_d, x_rest_para_d,
      x_rg_stat_d, x_ene_d;
  DevBuf<int> x_term_off_d, x_dcf_pairs_d, x_fr_pairs_d, x_rest_id_d, x_err_d;
  DevBuf<int2> x_terms_d;
  Params P{};
  Ptrs D{};
  sis_options opt{};
  sis_params prm{};
  int N = 0, R = 0;
  int parity = 0, precision = 0, mode = 1;
  uint64_t seed = 0;
  int64_t istep = 0, rebuilds = 0;
  bool lists_valid = false;
  // the lists follow the rules of the DCD / CHECK_FORCE jobs' list_bp,
  // list_exv and list_ele (P.list_rule = 1 at their build), not those of
  // the MD job's neighbour list; a step rebuilds them first
  bool lists_analysis = false;
  bool use_cells = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms[4] = {0, 0, 0, 0};
  // HIP events around each launch of the dominant kernel of a step call
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  double dom_ms = 0.0;
  int64_t dom_launches = 0;
  void tmark() {
    if (tev_used >= tev.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      tev.push_back(e);
    }
    HIPCHK(hipEventRecord(tev[tev_used++], stream));
  }
  void tcollect() {
    dom_ms = 0.0;
    dom_launches = (int64_t)(tev_used / 2);
    for (size_t k = 0; k + 1 < tev_used; k += 2) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, tev[k], tev[k + 1]));
      dom_ms += ms;
    }
    tev_used = 0;
  }
  // host copies
  std::vector<Topo> topo;
  std::vector<double> tK, ionic, mass, fric;
  std::vector<int> rep2lab, lab2rep;
  std::vector<double> lab2val;
  std::vector<double> lab2ion;   // ionic strength of each label (REMD over the ionic strength)
  std::vector<double> lab2rg;    // Rg-bias target of each label (REMD over the Rg target)
  bool rg_remd = false;          // the Rg targets travel with the labels (sis_gpu_set_rep_rg)
  std::vector<double> lab2dcf;   // constant-force magnitude of each label (REMD over the force)
  double dcf_dir[3] = {0, 0, 0}; // unit direction of the one constant-force pair
  bool dcf_remd = false;
  int idx_pair = 0;
  int pair_sets = 2;         // 2 x the number of replica dimensions
  std::vector<int> pair_tb;  // [pair_sets][R+2]
  std::vector<uint32_t> rep_tape;
  int64_t rep_tape_pos = 0;
  std::vector<int64_t> tape_off_h;
  BpPara bpp_h[7] = {};   // base-pair types 1..6 (GC AU GU CG UA UG)
  // device buffers
  DevBuf<Topo> topo_d;
  DevBuf<BpPara> bpp_d;
  DevBuf<int> bp3_hash_d;
  DevBuf<double> bp3_dH_d, bp3_dS_d, md_a_d, md_cf_d, md_cx_d, mass_d, md_cr_d;
  DevBuf<double> tK_d, ele_coef_d, lambdaD_d, ele_cut_d, box_d;
  DevBuf<double4> pos_d, pos2_d, vel_d, acc_d, move_d, frc_d, fbond_d;
  DevBuf<int> nb_ele_d, n_ele_d, nb_wca_d, n_wca_d, bp_tmp_d, n_bp_tmp_d, bp_off_d, bp_mp_d,
      n_bp_d, inc_d, n_inc_d, nt_cnt_d, comp_d, flags_d;
  DevBuf<double> bp_coef_d, ene_bp_d, for_bp_d;
  DevBuf<uint8_t> bp_status_d, bp_status_mc_d, bp_status_keep_d;
  DevBuf<int> cell_of_d, cell_cnt_d, cell_start_d, cell_beads_d, cell_fill_d;
  DevBuf<double> blk_top2_d;
  DevBuf<int> rebuild_flag_d, mask_d, mc_d, scan_bsum_d, scan_total_d;
  DevBuf<int> exc_cnt_d, exc_off_d, exc_fill_d, exc_list_d, ntlist_d, bpseq_d, seg_d;
  DevBuf<int> huge_d;   // per replica: a conflict component above the bucket capacity this step
  DevBuf<double> kin_d; // per replica kinetic energy (sis_gpu_kinetic)
  DevBuf<uint32_t> tape_d;
  DevBuf<int64_t> tape_off_d, tape_pos_d;
  DevBuf<double> noise_d, bm_cache_d, ene_part_d, ene_out_d, ecT_d, eCoef_d, eLam_d;
  DevBuf<unsigned long long> pers_bar_d;
  int batch_capb = 0;
  DevBuf<int> bb_mp_d, bb_done_d, bb_nbp_d, bb_rebuild_d, bb_todo_d, bb_off_d, bb_mcfirst_d, bb_nreb_d;
  DevBuf<int64_t> bb_stepno_d;
  DevBuf<double> bb_coef_d, bb_ene_d, bb_for_d;
  DevBuf<uint8_t> bb_status_d, bb_status_mc_d;
  DevBuf<int> pers_sync_d, chain_id_d, bp_val_d, bp_val2_d, bp_mp2_d;
  DevBuf<unsigned long long> bp_key_d, bp_key2_d;
  DevBuf<double> bp_coef2_d;
  DevBuf<char> sort_tmp_d;
  int nbp_total = 0;
  std::vector<int64_t> bp_rep_off;   // candidates of replica r: [off[r], off[r+1])
  std::vector<int> rebuilt;          // replicas the last rebuild() rebuilt
  int nblk_force = 0;
  int block = 256;
  int dev = 0;   // the HIP device current when the context was created
  // replayed per-phase steps (phases_graph_steps)
  DevBuf<int64_t> g_step_d;
  DevBuf<int> g_int_d;
  // device exchange of the batched form (sis_exchange.h)
  DevBuf<int> dx_rep2lab_d, dx_lab2rep_d, dx_pair_tb_d, dx_idx_d, dx_has_d, dx_vsdiv_d, dx_nexch_d, dx_acc_d;
  DevBuf<double> dx_labT_d, dx_labCoef_d, dx_labLam_d, dx_labCr_d, dx_oppT_d, dx_oppCoef_d, dx_oppLam_d,
      dx_energy_d, dx_vsf_d;
  hipGraphExec_t dx_graph = nullptr;
  // the per-phase form's captured steps (phases_graph_steps), one graph
  // for each of the two position buffers a replay can start from
  hipGraphExec_t pg_exec[2] = {nullptr, nullptr};
  bool pg_eager = false;   // the same batches of steps issued directly instead (the autotuner's choice)
  PhaseGraphKey pg_key[2]{};
  int dx_graph_key[4] = {0, 0, 0, 0};   // block, precision, exchange cadence, dynamic LDS of the captured graph
  int64_t dev_exchanges = 0;            // exchanges done on the device

  Ptrs ptrs() {
    Ptrs d{};
    d.topo = topo_d.p;
    d.chain_id = chain_id_d.p;
    d.bp_key = bp_key_d.p;
    d.bp_val = bp_val_d.p;
    d.bp3_hash = bp3_hash_d.p;
    d.bp3_dH = bp3_dH_d.p;
    d.bp3_dS = bp3_dS_d.p;
    d.md_a = md_a_d.p; d.md_cf = md_cf_d.p; d.md_cx = md_cx_d.p; d.mass = mass_d.p;
    d.md_cr = md_cr_d.p;
    d.tK = tK_d.p; d.ele_coef = ele_coef_d.p; d.lambdaD = lambdaD_d.p; d.ele_cut = ele_cut_d.p;
    d.box = box_d.p;
    d.bpp = bpp_d.p;
    d.pos = pos_d.p; d.pos_out = pos2_d.p; d.fbond = fbond_d.p; d.vel = vel_d.p; d.acc = acc_d.p; d.move = move_d.p; d.frc = frc_d.p;
    d.maxnb = D.maxnb; d.maxw = D.maxw; d.maxbp = D.maxbp; d.maxinc = D.maxinc; d.maxbpb = D.maxbpb;
    d.nb_ele = nb_ele_d.p; d.n_ele = n_ele_d.p; d.nb_wca = nb_wca_d.p; d.n_wca = n_wca_d.p;
    d.bp_tmp = bp_tmp_d.p; d.n_bp_tmp = n_bp_tmp_d.p; d.bp_off = bp_off_d.p;
    d.bp_mp = bp_mp_d.p; d.bp_coef = bp_coef_d.p; d.n_bp = n_bp_d.p;
    d.inc = inc_d.p; d.n_inc = n_inc_d.p;
    d.bp_status = bp_status_d.p; d.bp_status_mc = bp_status_mc_d.p;
    d.ene_bp = ene_bp_d.p; d.for_bp = for_bp_d.p; d.nt_cnt = nt_cnt_d.p; d.comp = comp_d.p;
    d.flags = flags_d.p;
    d.cell_of = cell_of_d.p; d.cell_cnt = cell_cnt_d.p; d.cell_start = cell_start_d.p;
    d.cell_beads = cell_beads_d.p;
    d.ncell_x = D.ncell_x; d.ncell_y = D.ncell_y; d.ncell_z = D.ncell_z; d.ncell = D.ncell;
    d.blk_top2 = blk_top2_d.p; d.rebuild_flag = rebuild_flag_d.p;
    d.tape = tape_d.p; d.tape_off = tape_off_d.p; d.tape_pos = tape_pos_d.p;
    d.noise = noise_d.p; d.bm_cache = bm_cache_d.p;
    d.ene_acc = ene_out_d.p;
    d.ss_off = ss_on ? ss_off_d.p : nullptr;
    d.ss_partner = ss_on ? ss_partner_d.p : nullptr;
    return d;
  }
  ~sis_gpu_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    DevBuf<int> *ib[] = {&bp3_hash_d, &nb_ele_d, &n_ele_d, &nb_wca_d, &n_wca_d, &bp_tmp_d,
                         &n_bp_tmp_d, &bp_off_d, &bp_mp_d, &n_bp_d, &inc_d, &n_inc_d, &nt_cnt_d,
                         &comp_d, &flags_d, &cell_of_d, &cell_cnt_d, &cell_start_d,
                         &cell_beads_d, &cell_fill_d, &rebuild_flag_d, &mask_d, &mc_d,
                         &scan_bsum_d, &scan_total_d, &exc_cnt_d, &exc_off_d, &exc_fill_d,
                         &exc_list_d, &ntlist_d, &bpseq_d, &seg_d, &huge_d, &pers_sync_d, &chain_id_d,
                         &bp_val_d, &bp_val2_d, &bp_mp2_d};
    for (auto *b : ib) b->free_();
    DevBuf<double> *db[] = {&kin_d, &bp3_dH_d, &bp3_dS_d, &md_a_d, &md_cf_d, &md_cx_d, &mass_d,
                            &md_cr_d, &tK_d, &ele_coef_d, &lambdaD_d, &ele_cut_d, &box_d,
                            &bp_coef_d, &ene_bp_d, &for_bp_d, &blk_top2_d, &noise_d,
                            &bm_cache_d, &ene_part_d, &ene_out_d, &ecT_d, &eCoef_d, &eLam_d,
                            &bp_coef2_d};
    for (auto *b : db) b->free_();
    bpp_d.free_();
    DevBuf<double> *xb[] = {&x_rg0_d, &x_rg0_opp_d, &x_dcf_opp_d, &x_dcf_f_d, &x_fr_k_d, &x_fr_init_d, &x_fr_velo_d, &x_rest_rcut_d,
                            &x_rest_para_d, &x_rg_stat_d, &x_ene_d};
    for (auto *b : xb) b->free_();
    DevBuf<int> *xi[] = {&x_term_off_d, &x_dcf_pairs_d, &x_fr_pairs_d, &x_rest_id_d, &x_err_d, &ss_off_d, &ss_partner_d};
    for (auto *b : xi) b->free_();
    x_terms_d.free_();
    DevBuf<double4> *vb[] = {&pos_d, &pos2_d, &vel_d, &acc_d, &move_d, &frc_d, &fbond_d};
    for (auto *b : vb) b->free_();
    topo_d.free_();
    bp_status_d.free_();
    bp_status_mc_d.free_(); bp_status_keep_d.free_();
    tape_d.free_();
    tape_off_d.free_();
    tape_pos_d.free_();
    bp_key_d.free_();
    pers_bar_d.free_();
    bb_mp_d.free_(); bb_done_d.free_(); bb_nbp_d.free_(); bb_rebuild_d.free_(); bb_todo_d.free_();
    bb_off_d.free_(); bb_mcfirst_d.free_(); bb_nreb_d.free_();
    bb_stepno_d.free_(); bb_coef_d.free_(); bb_ene_d.free_(); bb_for_d.free_();
    bb_status_d.free_(); bb_status_mc_d.free_();
    bp_key2_d.free_();
    sort_tmp_d.free_();
    DevBuf<int> *dxi[] = {&dx_rep2lab_d, &dx_lab2rep_d, &dx_pair_tb_d, &dx_idx_d, &dx_has_d, &dx_vsdiv_d,
                          &dx_nexch_d, &dx_acc_d};
    for (auto *b : dxi) b->free_();
    DevBuf<double> *dxd[] = {&dx_labT_d, &dx_labCoef_d, &dx_labLam_d, &dx_labCr_d, &dx_oppT_d, &dx_oppCoef_d,
                             &dx_oppLam_d, &dx_energy_d, &dx_vsf_d};
    for (auto *b : dxd) b->free_();
    if (dx_graph) (void)hipGraphExecDestroy(dx_graph);
    for (auto &e : pg_exec)
      if (e) (void)hipGraphExecDestroy(e);
    g_step_d.free_();
    g_int_d.free_();
    for (auto e : tev) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// ---------------------------------------------------------------------
// Host-side set-up restating init_bp's triplet table
// (src/init_bp.F90:39-172) and BP cutoffs (src/init_bp.F90:321-346).
bool host_comp(int a, int b) {
  return (a == SQ_A && b == SQ_U) || (a == SQ_U && (b == SQ_A || b == SQ_G)) ||
         (a == SQ_G && (b == SQ_C || b == SQ_U)) || (a == SQ_C && b == SQ_G);
}
int host_hkey(int w, int x, int u, int z, int y, int v) {
  if (x < y) return 1 + w + x * 4 + u * 16 + z * 64 + y * 256 + v * 1024;
  return 1 + v + y * 4 + z * 16 + u * 64 + x * 256 + w * 1024;
}
// NN stack index (1..21) of 5'-w x-3' / 3'-z y-5' (src/const_idx.F90:230-330)
int nn_index(int w, int x, int z, int y) {
  static const char ch[5] = {'A', 'U', 'G', 'C', 'D'};
  static const char *tbl[21][2] = {
      {"GC_CG", nullptr}, {"CC_GG", "GG_CC"}, {"GA_CU", "UC_AG"}, {"CG_GC", nullptr},
      {"AC_UG", "GU_CA"}, {"CA_GU", "UG_AC"}, {"AG_UC", "CU_GA"}, {"UA_AU", nullptr},
      {"AU_UA", nullptr}, {"AA_UU", "UU_AA"}, {"GC_UG", "GU_CG"}, {"CU_GG", "GG_UC"},
      {"GG_CU", "UC_GG"}, {"CG_GU", "UG_GC"}, {"AU_UG", "GU_UA"}, {"GA_UU", "UU_AG"},
      {"UG_GU", nullptr}, {"UA_GU", "UG_AU"}, {"GG_UU", "UU_GG"}, {"GU_UG", nullptr},
      {"AG_UU", "UU_GA"}};
  char k[6] = {ch[w], ch[x], '_', ch[z], ch[y], 0};
  for (int i = 0; i < 21; i++)
    for (int a = 0; a < 2; a++)
      if (tbl[i][a] && std::strcmp(tbl[i][a], k) == 0) return i + 1;
  return 0;
}
void build_triplet_table(const sis_params &p, std::vector<int> &hash, std::vector<double> &dH,
                         std::vector<double> &dS) {
  hash.assign(4097, 0);
  dH.assign(469, 0.0);
  dS.assign(469, 0.0);
  auto endc = [&](int h, int cx, int cy, int ox, int oy) {
    bool au = (cx == SQ_A && cy == SQ_U) || (cx == SQ_U && cy == SQ_A);
    bool gu = (cx == SQ_G && cy == SQ_U) || (cx == SQ_U && cy == SQ_G);
    bool oau = (ox == SQ_A && oy == SQ_U) || (ox == SQ_U && oy == SQ_A);
    bool ogc = (ox == SQ_G && oy == SQ_C) || (ox == SQ_C && oy == SQ_G);
    bool ogu = (ox == SQ_G && oy == SQ_U) || (ox == SQ_U && oy == SQ_G);
    int idx = 0;
    if (au) idx = oau ? 1 : ogc ? 2 : ogu ? 3 : 0;
    else if (gu) idx = oau ? 5 : ogc ? 4 : ogu ? 6 : 0;
    if (idx) {
      dH[h] = dH[h] + p.nnend_dH[idx - 1] - p.bp_dHend0;
      dS[h] = dS[h] + (p.nnend_dS[idx - 1] - p.bp_dSend0) * 1.0e-3;
    }
  };
  int h = 0;
  for (int x = 0; x <= 2; x++)
    for (int y = x + 1; y <= 3; y++) {
      if (!host_comp(x, y)) continue;
      for (int w = 0; w <= 3; w++)
        for (int z = 0; z <= 3; z++) {
          bool cwz = host_comp(w, z);
          double h1 = 0.0, s1 = 0.0;
          if (cwz) {
            int q = nn_index(w, x, z, y);
            h1 = 0.5 * (p.nn_dH[q - 1] - p.bp_dH0);
            s1 = 0.5 * (p.nn_dS[q - 1] - p.bp_dS0);
          }
          for (int u = 0; u <= 3; u++)
            for (int v = 0; v <= 3; v++) {
              bool cuv = host_comp(u, v);
              if (!cwz && !cuv) continue;
              h++;
              hash[host_hkey(w, x, u, z, y, v)] = h;
              if (cuv) {
                int q = nn_index(x, u, y, v);
                dH[h] = h1 + 0.5 * (p.nn_dH[q - 1] - p.bp_dH0);
                dS[h] = (s1 + 0.5 * (p.nn_dS[q - 1] - p.bp_dS0)) * 1.0e-3;
              } else {
                dH[h] = h1;
                dS[h] = s1 * 1.0e-3;
              }
              if (p.nnend) {
                if (cwz && !cuv) endc(h, x, y, w, z);
                else if (!cwz && cuv) endc(h, x, y, u, v);
              }
            }
        }
    }
}

// set_ele (src/set_ele.F90:45-99)
void host_set_ele(const sis_params &p, double T, double I, double &coef, double &lam, double &cut) {
  const double MM_A = 87.740, MM_B = -0.40008, MM_C = 9.398e-4, MM_D = -1.410e-6;
  const double EPS0 = 8.8541878128e-12, ELE = 1.602176634e-19, BOLTZ_J = 1.380649e-23,
               N_AVO = 6.02214076e23;
  const double J2K = (1.0 / 4184.0) * N_AVO;
  double Tc = T - 273.15;
  double diele = MM_A + MM_B * Tc + MM_C * Tc * Tc + MM_D * Tc * Tc * Tc;
  double lb = ELE * ELE / (4.0 * PI_D * EPS0 * diele * BOLTZ_J * T) * 1.0e10;
  double Zp = p.length_per_charge / lb;
  coef = J2K * 1.0e10 * (ELE * ELE) / (4.0 * PI_D * EPS0 * diele);
  coef = coef * (Zp * Zp);
  lam = 1.0e10 * std::sqrt((1.0e-3 * EPS0 * diele * BOLTZ_J) / (2.0 * N_AVO * (ELE * ELE))) *
        std::sqrt(T / I);
  cut = (p.ele_cutoff_type == 1) ? p.ele_cutoff : p.ele_cutoff * lam;
}

double md_mass(int sq) {  // src/job_md.F90:116-133
  switch (sq) {
    case SQ_A: return 328.212;
    case SQ_G: return 344.212;
    case SQ_C: return 304.182;
    case SQ_U: return 305.164;
    default: return 320.443;
  }
}

inline dim3 grid_beads(int N, int R, int B) { return dim3((N + B - 1) / B, R); }

void scan_exclusive(sis_gpu_ctx *c, const int *in, int *out, int64_t n, int *total_dev) {
  const int B = 1024;
  int64_t nb = (n + B - 1) / B;
  if ((int64_t)c->scan_bsum_d.n < nb + 1) c->scan_bsum_d.alloc(nb + 1);
  hipLaunchKernelGGL(k_scan_blocks<B>, dim3(nb), dim3(B), 0, c->stream, in, out, c->scan_bsum_d.p, n);
  hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1), 0, c->stream, c->scan_bsum_d.p, (int)nb,
                     c->scan_total_d.p, out + n);
  hipLaunchKernelGGL(k_scan_add<B>, dim3(nb), dim3(B), 0, c->stream, out, c->scan_bsum_d.p, n);
  (void)total_dev;
}

void alloc_bp(sis_gpu_ctx *c);
void alloc_lists(sis_gpu_ctx *c) {
  size_t RN = (size_t)c->R * c->N;
  c->nb_ele_d.alloc(RN * c->D.maxnb);
  c->nb_wca_d.alloc(RN * c->D.maxw);
  c->bp_tmp_d.alloc(RN * c->D.maxbpb);
  c->inc_d.alloc(RN * c->D.maxinc);
  alloc_bp(c);
}
// base-pair candidate arrays of the packed layout (contents are lost)
void alloc_bp(sis_gpu_ctx *c) {
  c->bp_mp_d.alloc((size_t)c->D.maxbp * 8);
  c->bp_coef_d.alloc((size_t)c->D.maxbp * 2);
  c->bp_status_d.alloc(c->D.maxbp);
  c->bp_status_mc_d.alloc(c->D.maxbp);
  c->ene_bp_d.alloc(c->D.maxbp);
  c->for_bp_d.alloc((size_t)c->D.maxbp * 18);
  c->bpseq_d.alloc((size_t)c->D.maxbp + c->R + 2);
  c->bp_key_d.alloc(c->D.maxbp);
  c->bp_key2_d.alloc(c->D.maxbp);
  c->bp_val_d.alloc(c->D.maxbp);
  c->bp_val2_d.alloc(c->D.maxbp);
  c->bp_mp2_d.alloc((size_t)c->D.maxbp * 8);
  c->bp_coef2_d.alloc((size_t)c->D.maxbp * 2);
}

void setup_cells(sis_gpu_ctx *c) {
  c->use_cells = false;
  if (!c->prm.pbc) return;
  double rcut = std::max({c->P.wca_sigma, c->P.bp_cutoff_dist, 0.0}) + c->P.nl_margin;
  for (int r = 0; r < c->R; r++) {
    double ec = 0.0, lam, coef;
    host_set_ele(c->prm, c->anneal_on ? c->ele_T[r] : c->tK[r], c->ionic[r], coef, lam, ec);
    if (c->prm.ele_on) rcut = std::max(rcut, ec + c->P.nl_margin);
  }
  int nc[3];
  for (int k = 0; k < 3; k++) nc[k] = (int)std::floor(c->prm.box[k] / rcut);
  if (nc[0] < 3 || nc[1] < 3 || nc[2] < 3 || c->N < 1024) return;
  c->use_cells = true;
  c->D.ncell_x = nc[0];
  c->D.ncell_y = nc[1];
  c->D.ncell_z = nc[2];
  c->D.ncell = nc[0] * nc[1] * nc[2];
  size_t RC = (size_t)c->R * c->D.ncell;
  c->cell_of_d.alloc((size_t)c->R * c->N);
  c->cell_cnt_d.alloc(RC);
  c->cell_start_d.alloc(RC + 1);
  c->cell_fill_d.alloc(RC);
  c->cell_beads_d.alloc((size_t)c->R * c->N);
}

// Per-replica conditions: set_ele and set_md_coef (src/job_md.F90:700-767).
// set_cut: also the electrostatic cutoff, as init_ele sets it at set-up
// (src/init_ele.F90:73-101).  After an exchange set_ele changes lambdaD and
// the coefficient only; the cutoff follows at the replica's next rebuild
// (src/job_md.F90:556, src/neighbor_list.F90:80-82).
void upload_conditions(sis_gpu_ctx *c, bool set_cut = true) {
  int R = c->R, N = c->N;
  std::vector<double> coef(R), lam(R), cut(R), cr((size_t)R * N), box(3 * R);
  for (int r = 0; r < R; r++) {
    host_set_ele(c->prm, c->anneal_on ? c->ele_T[r] : c->tK[r], c->ionic[r], coef[r], lam[r], cut[r]);
    for (int i = 0; i < N; i++) {
      if (c->P.bd) {  // src/job_md.F90:755-761
        cr[(size_t)r * N + i] = std::sqrt(2.0 * KB_KCAL * c->tK[r] * c->prm.dt / c->fric[i]);
        continue;
      }
      double c1 = 0.5 * c->prm.dt * c->fric[i] / c->mass[i];
      double c2 = std::sqrt(1.0 / (1.0 + c1));
      cr[(size_t)r * N + i] =
          0.5 * c2 / c->mass[i] * std::sqrt(2.0 * c->fric[i] * KB_KCAL * c->tK[r] * c->prm.dt);
    }
    for (int k = 0; k < 3; k++) box[3 * r + k] = c->prm.box[k];
  }
  HIPCHK(hipMemcpyAsync(c->tK_d.p, c->tK.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->ele_coef_d.p, coef.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->lambdaD_d.p, lam.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (set_cut)
    HIPCHK(hipMemcpyAsync(c->ele_cut_d.p, cut.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->md_cr_d.p, cr.data(), cr.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->box_d.p, box.data(), box.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

// pbc_resize (src/pbc.F90:79-138): each axis moves to original + change x
// count, clamped at the target, until every axis has reached it
void pbc_resize_host(sis_gpu_ctx *c) {
  const double SMALL = 2.220446049250313e-16;   // epsilon(PI), src/const_phys.F90:15
  bool reach[3] = {false, false, false};
  for (int i = 0; i < 3; i++) {
    if (std::fabs(c->rz_change[i]) < SMALL) { reach[i] = true; continue; }
    if (c->rz_has_target && std::fabs(c->prm.box[i] - c->rz_target[i]) < SMALL) { reach[i] = true; continue; }
    c->rz_count[i]++;
    const double x = c->rz_orig[i] + c->rz_change[i] * (double)c->rz_count[i];
    if (c->rz_change[i] < 0.0) {
      if (c->rz_has_target && x - c->rz_target[i] < SMALL) { c->prm.box[i] = c->rz_target[i]; reach[i] = true; }
      else c->prm.box[i] = x;
    } else {
      if (c->rz_has_target && x - c->rz_target[i] > SMALL) { c->prm.box[i] = c->rz_target[i]; reach[i] = true; }
      else c->prm.box[i] = x;
    }
  }
  if (reach[0] && reach[1] && reach[2]) c->rz_on = false;
}

void set_mask(sis_gpu_ctx *c, const std::vector<int> &m) {
  HIPCHK(hipMemcpyAsync(c->mask_d.p, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
}
void set_mc(sis_gpu_ctx *c, const std::vector<int> &m) {
  HIPCHK(hipMemcpyAsync(c->mc_d.p, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
}

// neighbor_list for the replicas in `mask` (src/neighbor_list.F90).
void rebuild(sis_gpu_ctx *c, const std::vector<int> &mask_in) {
  const int N = c->N, R = c->R, B = 256;
  std::vector<int> mask = mask_in;
  // Replicas that are not rebuilt keep their candidates, but the packed
  // array moves them: keep their limit-MC state to carry it over.
  const bool carry = R > 1 && c->P.nstep_bp_mc > 1 && (int64_t)c->bp_rep_off.size() == R + 1 &&
                     c->nbp_total > 0;
  if (carry) {
    if ((int64_t)c->bp_status_keep_d.n < c->nbp_total) c->bp_status_keep_d.alloc(c->nbp_total);
    HIPCHK(hipMemcpyAsync(c->bp_status_keep_d.p, c->bp_status_mc_d.p, c->nbp_total, hipMemcpyDeviceToDevice,
                          c->stream));
  }
  const std::vector<int64_t> old_off = c->bp_rep_off;
  set_mask(c, mask);
  for (int attempt = 0; attempt < 8; attempt++) {
    HIPCHK(hipMemsetAsync(c->flags_d.p, 0, 8 * sizeof(int), c->stream));
    Ptrs D = c->ptrs();
    hipLaunchKernelGGL(k_wrap, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, c->mask_d.p);
    if (c->use_cells) {
      size_t RC = (size_t)R * D.ncell;
      HIPCHK(hipMemsetAsync(c->cell_cnt_d.p, 0, RC * sizeof(int), c->stream));
      HIPCHK(hipMemsetAsync(c->cell_fill_d.p, 0, RC * sizeof(int), c->stream));
      hipLaunchKernelGGL(k_cell_count, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, c->mask_d.p);
      scan_exclusive(c, c->cell_cnt_d.p, c->cell_start_d.p, (int64_t)RC, nullptr);
      hipLaunchKernelGGL(k_cell_fill, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, c->mask_d.p,
                         c->cell_fill_d.p);
      hipLaunchKernelGGL(k_cell_sort, dim3((RC + B - 1) / B), dim3(B), 0, c->stream, c->P, D, c->mask_d.p);
      hipLaunchKernelGGL(k_search_cell, grid_beads(N, R, 128), dim3(128), 0, c->stream, c->P, D, c->mask_d.p);
    } else {
      hipLaunchKernelGGL(k_search_brute<128>, grid_beads(N, R, 128), dim3(128), 0, c->stream, c->P, D,
                         c->mask_d.p);
    }
    // BP candidate list in reference order over all replicas
    scan_exclusive(c, c->n_bp_tmp_d.p, c->bp_off_d.p, (int64_t)R * N, nullptr);
    HIPCHK(hipMemsetAsync(c->n_inc_d.p, 0, (size_t)R * N * sizeof(int), c->stream));
    hipLaunchKernelGGL(k_bp_fill, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, nullptr);
    int hflags[8];
    int total = 0;
    HIPCHK(hipMemcpyAsync(&total, c->bp_off_d.p + (size_t)R * N, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (total > 0 && total <= c->D.maxbp) {
      // candidates into the reference's list order, then their incidences
      size_t tb = 0;
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, c->bp_key_d.p, c->bp_key2_d.p, c->bp_val_d.p,
                                                c->bp_val2_d.p, total, 0, 64, c->stream));
      if (c->sort_tmp_d.n < tb) c->sort_tmp_d.alloc(tb);
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->sort_tmp_d.p, tb, c->bp_key_d.p, c->bp_key2_d.p, c->bp_val_d.p,
                                                c->bp_val2_d.p, total, 0, 64, c->stream));
      hipLaunchKernelGGL(k_bp_permute, dim3((total + B - 1) / B), dim3(B), 0, c->stream, D, total, c->bp_val2_d.p,
                         c->bp_mp2_d.p, c->bp_coef2_d.p);
      // back into the same buffers: captured step graphs keep their pointers
      HIPCHK(hipMemcpyAsync(c->bp_mp_d.p, c->bp_mp2_d.p, (size_t)total * 8 * sizeof(int), hipMemcpyDeviceToDevice,
                            c->stream));
      HIPCHK(hipMemcpyAsync(c->bp_coef_d.p, c->bp_coef2_d.p, (size_t)total * 2 * sizeof(double),
                            hipMemcpyDeviceToDevice, c->stream));
      hipLaunchKernelGGL(k_bp_inc, dim3((total + B - 1) / B), dim3(B), 0, c->stream, c->P, D, total);
    }
    hipLaunchKernelGGL(k_inc_sort, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, nullptr);
    HIPCHK(hipMemcpyAsync(hflags, c->flags_d.p, sizeof hflags, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    bool grow = false;
    if (hflags[2] > c->D.maxnb) { c->D.maxnb = hflags[2] + hflags[2] / 4 + 8; grow = true; }
    if (hflags[3] > c->D.maxw) { c->D.maxw = hflags[3] + hflags[3] / 4 + 4; grow = true; }
    if (hflags[4] > c->D.maxbpb) { c->D.maxbpb = hflags[4] + hflags[4] / 4 + 4; grow = true; }
    if (hflags[6] > c->D.maxinc) { c->D.maxinc = hflags[6] + hflags[6] / 4 + 8; grow = true; }
    if (total > c->D.maxbp) { c->D.maxbp = total + total / 4 + 64; grow = true; }
    if (!grow && hflags[0]) {  // overflow reported without sizes: double everything
      c->D.maxnb *= 2; c->D.maxw *= 2; c->D.maxbpb *= 2; c->D.maxinc *= 2; c->D.maxbp *= 2;
      grow = true;
    }
    if (!grow) {
      c->nbp_total = total;
      hipLaunchKernelGGL(k_clear_move, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, c->ptrs(),
                         c->mask_d.p);
      c->bp_rep_off.assign(R + 1, 0);
      if (R == 1) {
        c->bp_rep_off[1] = total;
      } else if (c->P.nstep_bp_mc > 1) {
        std::vector<int> off((size_t)R * N + 1);
        HIPCHK(hipMemcpyAsync(off.data(), c->bp_off_d.p, off.size() * sizeof(int), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int r = 0; r <= R; r++) c->bp_rep_off[r] = off[(size_t)r * N];
        if (carry)
          for (int r = 0; r < R; r++) {
            if (mask[r]) continue;
            int64_t n0 = old_off[r + 1] - old_off[r], n1 = c->bp_rep_off[r + 1] - c->bp_rep_off[r];
            if (n0 != n1) throw ApiError{SIS_ERR_STATE, "candidate count of a kept replica changed"};
            if (n1 > 0)
              HIPCHK(hipMemcpyAsync(c->bp_status_mc_d.p + c->bp_rep_off[r], c->bp_status_keep_d.p + old_off[r],
                                    n1, hipMemcpyDeviceToDevice, c->stream));
          }
      }
      c->rebuilt = mask;
      c->lists_valid = true;
      c->lists_analysis = c->P.list_rule != 0;
      return;
    }
    // lists of every replica are rebuilt after a resize
    std::fill(mask.begin(), mask.end(), 1);
    set_mask(c, mask);
    alloc_lists(c);
  }
  throw ApiError{SIS_ERR_CAPACITY, "neighbour lists did not converge in size"};
}

// The cutoff-free lists of the DCD and CHECK_FORCE jobs (src/main.F90:144-154:
// list_bp, list_exv, list_ele) for every replica at the positions held.
// Pairs beyond the neighbour cutoffs are left out: none of them has an
// energy, a force or a formed base pair, so only the rules of the three
// routines change (no circular wrap in the exclusions, list_ele's bond and
// angle exclusions on every chain pair, base-pair candidates with
// jmp >= imp + bp_min_loop + 1 over the system, in (imp, jmp) order).
void analysis_rebuild(sis_gpu_ctx *c) {
  c->P.list_rule = 1;
  c->lists_valid = false;
  try {
    rebuild(c, std::vector<int>(c->R, 1));
  } catch (...) {
    c->P.list_rule = 0;
    c->lists_valid = false;
    throw;
  }
  c->P.list_rule = 0;
}

// Base-pair evaluation + limit MC (force_bp_limit_triplet / force_bp_triplet).
// mc: limit-MC flags per replica, uploaded here; null when the flags are
// already on the device (replayed graph steps, whose step number is
// step_dev's)
void bp_phase(sis_gpu_ctx *c, const std::vector<int> *mc, const double *ecT_dev, const int *mask_dev,
              const int64_t *step_dev = nullptr, bool resets_done = false) {
  const int N = c->N, R = c->R, B = 256;
  if (mc) set_mc(c, *mc);
  Ptrs D = c->ptrs();
  bool any_mc = !mc;
  if (mc)
    for (int r = 0; r < R; r++) any_mc |= (*mc)[r] != 0;
  any_mc = any_mc && c->P.nstep_bp_mc > 0;
  // performance-form MC: counts, union-find and buckets reset in one pass,
  // the union done by the evaluation itself, the components with excess
  // listed for k_mc_list
  const bool perf_mc = any_mc && !c->parity;
  const size_t RN = (size_t)R * N;
  if (perf_mc) {
    if (c->seg_d.n < RN * SEGCAP) c->seg_d.alloc(RN * SEGCAP);
    if (c->exc_list_d.n < RN + 1) c->exc_list_d.alloc(RN + 1);
    if (!resets_done)
      hipLaunchKernelGGL(k_mc_reset, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, c->exc_cnt_d.p,
                         c->exc_fill_d.p, c->exc_list_d.p, mask_dev);
  } else if (!resets_done) {
    hipLaunchKernelGGL(k_reset_cnt, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, mask_dev);
  }
  // captured steps launch over the candidate capacity and read the count
  // on the device, so that their graph stays valid across rebuilds
  const bool dev_count = step_dev != nullptr;
  const int nbp_arg = dev_count ? -1 : (int)c->nbp_total;
  const int64_t nbp_grid = dev_count ? c->D.maxbp : c->nbp_total;
  if (nbp_grid > 0) {
    if (c->precision == 1)
      hipLaunchKernelGGL(k_bp_eval<float>, dim3((nbp_grid + B - 1) / B), dim3(B), 0, c->stream, c->P, D,
                         nbp_arg, c->mc_d.p, ecT_dev, mask_dev, perf_mc ? 1 : 0);
    else
      hipLaunchKernelGGL(k_bp_eval<double>, dim3((nbp_grid + B - 1) / B), dim3(B), 0, c->stream, c->P, D,
                         nbp_arg, c->mc_d.p, ecT_dev, mask_dev, perf_mc ? 1 : 0);
  }
  if (!any_mc) return;
  if (c->parity) {
    hipLaunchKernelGGL(k_mc_serial, dim3(R), dim3(64), 0, c->stream, c->P, D, c->mc_d.p, ecT_dev,
                       c->ntlist_d.p, c->bpseq_d.p, mask_dev);
  } else {
    if (nbp_grid == 0) return;
    hipLaunchKernelGGL(k_seg_fill, dim3((nbp_grid + B - 1) / B), dim3(B), 0, c->stream, c->P, D,
                       nbp_arg, c->exc_cnt_d.p, c->seg_d.p, c->exc_fill_d.p, c->mc_d.p, mask_dev,
                       c->exc_list_d.p);
    // one block per listed component as far as the device holds them at
    // once (C4 lists about 3,400 a step): a block that takes two in turn
    // can start the longest deletion chain late
    const int nlist = (int)std::min<int64_t>((int64_t)std::max<int64_t>(RN, 1), 8192);
    hipLaunchKernelGGL(k_mc_list, dim3(nlist), dim3(64), 0, c->stream, c->P, D, c->exc_cnt_d.p, c->seg_d.p,
                       c->exc_list_d.p, ecT_dev, c->seed, c->istep, c->huge_d.p, step_dev);
    hipLaunchKernelGGL(k_mc_huge, dim3(R), dim3(64), 0, c->stream, c->P, D, c->exc_cnt_d.p, c->seg_d.p,
                       c->exc_fill_d.p, c->huge_d.p, ecT_dev, c->seed, c->istep, step_dev);
  }
}

// launch shape of the pair/integrator kernel: FORCE_B threads per
// workgroup, FORCE_LPB lanes per bead; force_blocks = workgroups per replica
constexpr int FORCE_B = 256, FORCE_LPB = 8;
inline int force_blocks(int N) { return (N + FORCE_B / FORCE_LPB - 1) / (FORCE_B / FORCE_LPB); }

// mark: time the pair/integrator kernel with the context's event pairs
void force_kernel(sis_gpu_ctx *c, int integrate, const int *mask_dev, const int64_t *step_dev = nullptr,
                  bool mark = false) {
  const int N = c->N, R = c->R;
  constexpr int B = FORCE_B, LPB = FORCE_LPB;
  dim3 g(force_blocks(N), R);
  c->nblk_force = g.x;
  if ((int64_t)c->blk_top2_d.n < (int64_t)R * g.x * 2) c->blk_top2_d.alloc((size_t)R * g.x * 2);
  // bonded forces in the pair kernel, unless external terms add to them
  const int inl = c->ext_on ? 0 : 1;
  if (!inl) hipLaunchKernelGGL((k_bonded<LPB, B>), g, dim3(B), 0, c->stream, c->P, c->ptrs(), mask_dev);
  if (c->ext_on) {
    if (c->X.rg) hipLaunchKernelGGL(k_rg_stats<256>, dim3(R), dim3(256), 0, c->stream, c->P, c->ptrs(), c->X);
    hipLaunchKernelGGL(k_external<256>, grid_beads(N, R, 256), dim3(256), 0, c->stream, c->P, c->ptrs(), c->X,
                       c->istep, mask_dev);
  }
  if (mark) c->tmark();
  if (c->precision == 1)
    hipLaunchKernelGGL((k_force_lanes<float, LPB, B>), g, dim3(B), 0, c->stream, c->P, c->ptrs(), integrate,
                       c->parity, c->seed, c->istep, mask_dev, step_dev, inl);
  else
    hipLaunchKernelGGL((k_force_lanes<double, LPB, B>), g, dim3(B), 0, c->stream, c->P, c->ptrs(), integrate,
                       c->parity, c->seed, c->istep, mask_dev, step_dev, inl);
  if (mark) c->tmark();
}

void check_flags(sis_gpu_ctx *c) {
  int f[8];
  HIPCHK(hipMemcpyAsync(f, c->flags_d.p, sizeof f, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (f[7]) throw ApiError{SIS_ERR_STATE, "limit MC met a base-pair candidate outside the system"};
  if (f[1] == 2) throw ApiError{SIS_ERR_STATE, "limit MC found an excess nucleotide without pairs"};
  if (f[5]) throw ApiError{SIS_ERR_CAPACITY, "a base-pair conflict component outgrew the limit-MC scratch"};
  if (f[1]) throw ApiError{SIS_ERR_STATE, "parity random-number tape exhausted"};
}

// kinetic energy of every replica into e[R]
void kinetic_all(sis_gpu_ctx *c, double *e) {
  const int R = c->R;
  if (c->kin_d.n < (size_t)R) c->kin_d.alloc(R);
  hipLaunchKernelGGL(k_kinetic<256>, dim3(R), dim3(256), 0, c->stream, c->P, c->ptrs(), c->kin_d.p);
  HIPCHK(hipMemcpyAsync(e, c->kin_d.p, R * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

// energy_sumup for replica mask at conditions (ecT, coef, lam per replica)
void energy_phase(sis_gpu_ctx *c, const std::vector<int> &mask, const std::vector<double> &T,
                  int flags, double *out /* R*16 */, const std::vector<double> *ionic = nullptr,
                  const std::vector<double> *rg0 = nullptr, const std::vector<double> *dcf_k = nullptr) {
  const int N = c->N, R = c->R;
  const int B = 256;
  std::vector<double> coef(R), lam(R);
  for (int r = 0; r < R; r++) {
    double cut;
    // an annealed run keeps the electrostatics of its first temperature
    // (set_ele runs once, src/main.F90:90-97 before init_ele)
    host_set_ele(c->prm, c->anneal_on ? c->ele_T[r] : T[r], ionic ? (*ionic)[r] : c->ionic[r], coef[r], lam[r],
                 cut);
  }
  HIPCHK(hipMemcpyAsync(c->ecT_d.p, T.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->eCoef_d.p, coef.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->eLam_d.p, lam.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  set_mask(c, mask);
  if (!(flags & 2)) {
    std::vector<int> mc(R, 0);
    for (int r = 0; r < R; r++) mc[r] = (flags & 1) && mask[r];
    bp_phase(c, &mc, c->ecT_d.p, c->mask_d.p);
  }
  dim3 g = grid_beads(N, R, B);
  if ((int64_t)c->ene_part_d.n < (int64_t)R * g.x * 8) c->ene_part_d.alloc((size_t)R * g.x * 8);
  hipLaunchKernelGGL(k_energy<B>, g, dim3(B), 0, c->stream, c->P, c->ptrs(), c->ecT_d.p, c->eCoef_d.p,
                     c->eLam_d.p, c->ene_part_d.p, c->mask_d.p, flags & 1, -1, 0.0, 0.0);
  const double *ext = nullptr;
  if (c->ext_on) {
    if (c->X.rg) hipLaunchKernelGGL(k_rg_stats<256>, dim3(R), dim3(256), 0, c->stream, c->P, c->ptrs(), c->X);
    ExtDev X = c->X;
    if (rg0 && X.rg) {   // Rg targets of other labels (src/energy_replica.F90:107-109)
      if (c->x_rg0_opp_d.n < (size_t)R) c->x_rg0_opp_d.alloc(R);
      HIPCHK(hipMemcpyAsync(c->x_rg0_opp_d.p, rg0->data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
      X.rg0 = c->x_rg0_opp_d.p;
    }
    std::vector<double> fo;
    if (dcf_k && X.ndcf == 1) {   // forces of other labels (src/energy_replica.F90:103-105)
      fo.resize((size_t)3 * R);
      for (int r = 0; r < R; r++)
        for (int k = 0; k < 3; k++) fo[3 * r + k] = (*dcf_k)[r] * c->dcf_dir[k];
      if (c->x_dcf_opp_d.n < fo.size()) c->x_dcf_opp_d.alloc(fo.size());
      HIPCHK(hipMemcpyAsync(c->x_dcf_opp_d.p, fo.data(), fo.size() * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
      X.dcf_f = c->x_dcf_opp_d.p;
    }
    hipLaunchKernelGGL(k_external_energy<256>, dim3(R), dim3(256), 0, c->stream, c->P, c->ptrs(), X);
    ext = c->x_ene_d.p;
  }
  hipLaunchKernelGGL(k_energy_reduce, dim3(R), dim3(1), 0, c->stream, c->ene_part_d.p, (int)g.x, c->ene_out_d.p,
                     ext, -1);
  HIPCHK(hipMemcpyAsync(out, c->ene_out_d.p, (size_t)R * 16 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  check_flags(c);
  if (c->ext_on && c->X.stage) {
    int err = 0;
    HIPCHK(hipMemcpy(&err, c->x_err_d.p, sizeof err, hipMemcpyDeviceToHost));
    if (err) {
      HIPCHK(hipMemset(c->x_err_d.p, 0, sizeof err));
      // src/energy_stage.F90:23-27 stops the run here
      throw ApiError{SIS_ERR_STATE, "Z-coordinate < 0.85 * stage_sigma in energy_stage"};
    }
  }
}

int get_pair(sis_gpu_ctx *c, int lab) { return c->pair_tb[c->idx_pair * (c->R + 2) + lab]; }

// exchange_pair_tb (src/var_replica.F90:162-259) of a grid of replicas with
// dims[d] values along dimension d (R = their product; labels laid out as
// init_replica does, the first dimension slowest): set parity * ndim + d
// pairs, along dimension d, labels l and l + n (n = R / (dims[0] ...
// dims[d])) from the first value on (odd parity) or the second (even);
// 0 marks a label without a partner.  Sets are used in that order.
std::vector<int> make_pair_tb(int R, const std::vector<int> &dims) {
  const int nd = (int)dims.size();
  std::vector<int> tb((size_t)2 * nd * (R + 2), 0);
  for (int par = 0; par < 2; par++) {
    int ndiv = 1;
    for (int d = 0; d < nd; d++) {
      int *t = tb.data() + (size_t)(par * nd + d) * (R + 2);
      const int nv = dims[d], ndiv_pre = ndiv;
      ndiv *= nv;
      const int n = R / ndiv;
      int iset = 0;
      for (int cyc = 0; cyc < ndiv_pre; cyc++) {
        if (par == 1)
          for (int k = 0; k < n; k++) t[++iset] = 0;
        for (int v = par == 0 ? 1 : 2; v <= nv - 1; v += 2)
          for (int k = 1; k <= 2 * n; k++) {
            ++iset;
            t[iset] = k <= n ? iset + n : iset - n;
          }
        // the last value has no partner when the pairs leave it out
        if ((par == 0) == (nv % 2 == 1))
          for (int k = 0; k < n; k++) t[++iset] = 0;
      }
    }
  }
  return tb;
}

double rep_uniform(sis_gpu_ctx *c, uint64_t k) {
  if (c->parity) {
    if (c->rep_tape_pos + 2 > (int64_t)c->rep_tape.size())
      throw ApiError{SIS_ERR_STATE, "parity exchange tape exhausted"};
    uint32_t w1 = c->rep_tape[c->rep_tape_pos], w2 = c->rep_tape[c->rep_tape_pos + 1];
    c->rep_tape_pos += 2;
    double a = (double)(w1 >> 5), b = (double)(w2 >> 6);
    return (a * 67108864.0 + b) * (1.0 / 9007199254740991.0);
  }
  U4 o = philox(U4{(uint32_t)c->istep, (uint32_t)(c->istep >> 32), (uint32_t)k, 0x72657863u},
                (uint32_t)c->seed, (uint32_t)(c->seed >> 32) ^ 0x1234567u);
  uint64_t m = (((uint64_t)o.x << 21) ^ (uint64_t)(o.y >> 11)) & ((1ull << 53) - 1);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

void dcf_upload(sis_gpu_ctx *c);

// energy_replica + replica_exchange (src/energy_replica.F90:47-156,
// src/replica_exchange.F90:35-154) with the current set of the pair table.
void do_exchange(sis_gpu_ctx *c, double *replica_energies, int32_t *accepted, double *own_energies = nullptr) {
  const int R = c->R;
  std::vector<int> all(R, 1);
  if (!c->lists_valid) rebuild(c, all);
  std::vector<double> e1((size_t)R * 16), e2((size_t)R * 16);
  int mcflag = c->P.nstep_bp_mc > 0 ? 1 : 0;
  energy_phase(c, all, c->tK, mcflag, e1.data());
  // counterpart conditions; the BP state of the own-condition pass is kept
  // (temperature and ionic strength of the partner's label,
  // src/energy_replica.F90:90-113)
  std::vector<double> Topp(R), Iopp(R), Gopp(R), Fopp(R);
  std::vector<int> has(R, 0);
  for (int r = 0; r < R; r++) {
    int lab = c->rep2lab[r + 1], opp = get_pair(c, lab);
    if (opp > 0 && opp <= R) {
      Topp[r] = c->lab2val[opp]; Iopp[r] = c->lab2ion[opp]; Gopp[r] = c->lab2rg[opp]; Fopp[r] = c->lab2dcf[opp];
      has[r] = 1;
    } else {
      Topp[r] = c->tK[r]; Iopp[r] = c->ionic[r]; Gopp[r] = c->lab2rg[lab]; Fopp[r] = c->lab2dcf[lab];
    }
  }
  size_t nb = c->nbp_total;
  std::vector<uint8_t> st(nb);
  std::vector<double> en(nb);
  if (nb) {
    HIPCHK(hipMemcpyAsync(st.data(), c->bp_status_d.p, nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(en.data(), c->ene_bp_d.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  energy_phase(c, has, Topp, mcflag, e2.data(), &Iopp, c->rg_remd ? &Gopp : nullptr,
               c->dcf_remd ? &Fopp : nullptr);
  if (nb) {
    HIPCHK(hipMemcpyAsync(c->bp_status_d.p, st.data(), nb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ene_bp_d.p, en.data(), nb * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  if (own_energies) std::copy(e1.begin(), e1.end(), own_energies);
  std::vector<double> E(2 * (R + 1), 0.0);
  for (int r = 0; r < R; r++) {
    E[2 * (r + 1)] = e1[(size_t)r * 16];
    E[2 * (r + 1) + 1] = has[r] ? e2[(size_t)r * 16] : 0.0;
  }
  if (replica_energies)
    for (int r = 0; r < R; r++) { replica_energies[2 * r] = E[2 * (r + 1)]; replica_energies[2 * r + 1] = E[2 * (r + 1) + 1]; }
  std::vector<double> rnd(R + 2);
  int nr = 0;
  for (int i = 1; i <= R; i += 2) {   // uniform k (0-based) of the step for pair k+1
    const double u = rep_uniform(c, nr);
    rnd[++nr] = u;
  }
  nr = 0;
  int nacc = 0;
  for (int li = 1; li <= R; li++) {
    int lj = get_pair(c, li);
    if (lj < li) continue;
    nr++;
    int ri = c->lab2rep[li], rj = c->lab2rep[lj];
    double tm = c->lab2val[li], tn = c->lab2val[lj];
    double pim = E[2 * ri], pin = E[2 * ri + 1], pjn = E[2 * rj], pjm = E[2 * rj + 1];
    double delta = ((pjm - pim) / tm - (pjn - pin) / tn) / KB_KCAL;
    bool acc = (delta < 0.0) || (rnd[nr] <= std::exp(-delta));
    if (accepted) accepted[nr - 1] = acc ? 1 : 0;
    if (acc) {
      nacc++;
      c->rep2lab[ri] = lj; c->rep2lab[rj] = li;
      c->lab2rep[li] = rj; c->lab2rep[lj] = ri;
      double sf = std::sqrt(tn / tm);
      const int B = 256;
      hipLaunchKernelGGL(k_scale_vel, dim3((c->N + B - 1) / B), dim3(B), 0, c->stream, c->P, c->ptrs(), ri - 1, sf, 0);
      hipLaunchKernelGGL(k_scale_vel, dim3((c->N + B - 1) / B), dim3(B), 0, c->stream, c->P, c->ptrs(), rj - 1, sf, 1);
    }
  }
  c->idx_pair = (c->idx_pair + 1) % c->pair_sets;   // set_forward (src/var_replica.F90:104-117)
  for (int r = 0; r < R; r++) {
    c->tK[r] = c->lab2val[c->rep2lab[r + 1]];
    c->ionic[r] = c->lab2ion[c->rep2lab[r + 1]];
  }
  if (c->ext_on && c->X.rg && c->rg_remd) {   // src/job_md.F90:564-566
    std::vector<double> g(R);
    for (int r = 0; r < R; r++) g[r] = c->lab2rg[c->rep2lab[r + 1]];
    HIPCHK(hipMemcpyAsync(c->x_rg0_d.p, g.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (c->ext_on && c->dcf_remd) dcf_upload(c);   // src/job_md.F90:559-562
  // an exchange permutes the conditions, so the largest electrostatic
  // cutoff and the cell grid stay as they are; each replica keeps its
  // cutoff until its next rebuild
  upload_conditions(c, false);
  (void)nacc;
}

// One job_md step for every replica (src/job_md.F90:362-460), one launch
// per phase.
void dcd_set_box(sis_gpu_ctx *c);   // sis_output_host.h

// constant forces of every replica from its label's magnitude
void dcf_upload(sis_gpu_ctx *c) {
  const int R = c->R;
  std::vector<double> f((size_t)3 * R);
  for (int r = 0; r < R; r++)
    for (int k = 0; k < 3; k++) f[3 * r + k] = c->lab2dcf[c->rep2lab[r + 1]] * c->dcf_dir[k];
  HIPCHK(hipMemcpyAsync(c->x_dcf_f_d.p, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

// current entry of the timed Rg bias: k and the target of every replica
void rg_schedule_apply(sis_gpu_ctx *c) {
  const int n = (int)c->rgs_step.size();
  c->X.rg_k = c->rgs_k[c->irgs - 1];
  std::vector<double> r0(c->R, c->rgs_0[c->irgs - 1]);
  std::fill(c->lab2rg.begin(), c->lab2rg.end(), c->rgs_0[c->irgs - 1]);
  // on the engine's stream: the steps already queued read the old targets
  HIPCHK(hipMemcpyAsync(c->x_rg0_d.p, r0.data(), c->R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->rgs_next = c->irgs < n ? c->rgs_step[c->irgs] : 0;
}

// The resize block of src/job_md.F90:576-596: new box, then every
// replica's lists rebuilt and displacements zeroed; the next step's skin
// test starts from zero and does not force the limit MC
void box_resize_now(sis_gpu_ctx *c) {
  pbc_resize_host(c);
  dcd_set_box(c);
  upload_conditions(c);
  setup_cells(c);
  std::vector<int> all(c->R, 1);
  rebuild(c, all);
  HIPCHK(hipMemsetAsync(c->rebuild_flag_d.p, 0, c->R * sizeof(int), c->stream));
}

void step_phases(sis_gpu_ctx *c) {
  const int N = c->N, R = c->R;
  c->istep++;
  int mc_all = (c->P.nstep_bp_mc > 0 && c->istep % c->P.nstep_bp_mc == 0) ? 1 : 0;
  std::vector<int> flag(R, 0), mc(R, mc_all);
  if (!c->lists_valid) {
    std::fill(flag.begin(), flag.end(), 1);
  } else if (c->opt.rebuild_every > 0) {
    if ((c->istep - 1) % c->opt.rebuild_every == 0 && c->istep > 1) std::fill(flag.begin(), flag.end(), 1);
  } else {
    HIPCHK(hipMemcpyAsync(flag.data(), c->rebuild_flag_d.p, R * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  // A rebuild forces MC (src/job_md.F90:392-394).  The reference keeps one
  // flag for its replica loop, so replicas after a rebuilding one also do
  // MC in that step; parity mode follows it.  Otherwise the MC decision of
  // a replica depends on that replica alone, so that results do not depend
  // on which other replicas share the launch.
  bool any = false, sticky = false;
  for (int r = 0; r < R; r++) any = any || flag[r];
  if (any) { rebuild(c, flag); c->rebuilds++; flag = c->rebuilt; }
  for (int r = 0; r < R; r++) {
    if (flag[r]) sticky = c->parity != 0;
    if ((flag[r] || sticky) && c->P.nstep_bp_mc > 0) mc[r] = 1;
  }
  std::vector<int> all(R, 1);
  set_mask(c, all);
  bp_phase(c, &mc, nullptr, nullptr);
  if (c->P.nstep_bp_mc > 1) {
    // bp_status_MC = bp_status of the replicas that did MC
    // (src/job_md.F90:399-401); candidates are grouped by replica
    if (mc_all) {
      HIPCHK(hipMemcpyAsync(c->bp_status_mc_d.p, c->bp_status_d.p, c->nbp_total, hipMemcpyDeviceToDevice, c->stream));
    } else if (any) {
      for (int r = 0; r < R; r++) {
        if (!mc[r]) continue;
        int64_t b0 = c->bp_rep_off[r], b1 = c->bp_rep_off[r + 1];
        if (b1 > b0)
          HIPCHK(hipMemcpyAsync(c->bp_status_mc_d.p + b0, c->bp_status_d.p + b0, b1 - b0, hipMemcpyDeviceToDevice, c->stream));
      }
    }
  }
  if (c->parity)
    hipLaunchKernelGGL(k_noise_parity, dim3(1), dim3(1), 0, c->stream, c->P, c->ptrs(), 3 * N, 0, nullptr);
  force_kernel(c, 1, nullptr, nullptr, true);
  std::swap(c->pos_d, c->pos2_d);   // updated positions become current
  hipLaunchKernelGGL(k_skin<256>, dim3(R), dim3(256), 0, c->stream, c->P, c->ptrs(), c->nblk_force, GStep{});
}

// Whether the per-phase form can replay its steps from a captured graph:
// the performance form without external terms or a fixed rebuild cadence.
bool graph_steps_ok(sis_gpu_ctx *c) {
  return !c->parity && !c->ext_on && c->opt.rebuild_every == 0 && !(c->opt.flags & SIS_OPT_NO_GRAPH);
}

// The per-phase form's steps (step_phases) replayed from a HIP graph of
// GRAPH_STEPS steps, with no host round trip per step: the step number and
// the limit-MC flags come from k_gstep_begin, and k_skin halts the steps
// that follow when a replica needs a rebuild (or the requested count is
// reached).  The host then rebuilds, captures again (rebuilds resize the
// launches) and replays on.  Same kernels, same arithmetic and random
// streams as step_phases.
constexpr int GRAPH_STEPS = 16;   // even: the positions end in the buffer they started in

// Runs batches of GRAPH_STEPS steps (launch() issues one batch: a captured
// graph's replay or the same launches issued directly) until a replica
// halts them (a rebuild is due) or the requested step is reached.
template <typename F>
void run_phase_batches(sis_gpu_ctx *c, F &&launch) {
  int halt = 0;
  int64_t now = c->istep;
  while (!halt) {
    launch();
    HIPCHK(hipMemcpyAsync(&halt, c->g_int_d.p, sizeof halt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&now, c->g_step_d.p, sizeof now, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!halt && now - c->istep != (int64_t)GRAPH_STEPS) throw ApiError{SIS_ERR_STATE, "graph steps did not advance"};
    if (!halt) c->istep = now;
  }
  // an odd number of steps in the last replay leaves the positions in the
  // other buffer
  if ((now - c->istep) % 2) std::swap(c->pos_d, c->pos2_d);
  c->istep = now;
}

void phases_graph_steps(sis_gpu_ctx *c, int64_t nsteps) {
  const int R = c->R;
  if (c->g_step_d.n < 3) c->g_step_d.alloc(3);
  if (c->g_int_d.n < (size_t)(1 + 3 * R)) {
    c->g_int_d.alloc(1 + 3 * R);
    HIPCHK(hipMemset(c->g_int_d.p, 0, (1 + 3 * (size_t)R) * sizeof(int)));
  }
  GStep G{};
  G.istep = c->g_step_d.p;
  G.istep_cur = c->g_step_d.p + 1;
  G.stop = c->g_step_d.p + 2;
  G.halt = c->g_int_d.p;
  G.active = c->g_int_d.p + 1;
  G.force_mc = c->g_int_d.p + 1 + R;
  G.mc = c->mc_d.p;
  std::vector<int> fmc(R, 0), flag(R, 0);
  const int64_t stop = c->istep + nsteps;
  while (c->istep < stop) {
    // a rebuild due from the last skin test, or lists not built yet
    // (src/job_md.F90:376-395); the rebuilt replicas do the limit MC first
    if (!c->lists_valid) {
      std::fill(flag.begin(), flag.end(), 1);
    } else {
      HIPCHK(hipMemcpy(flag.data(), c->rebuild_flag_d.p, R * sizeof(int), hipMemcpyDeviceToHost));
    }
    bool any = false;
    for (int r = 0; r < R; r++) any = any || flag[r];
    if (any) {
      rebuild(c, flag);
      c->rebuilds++;
      for (int r = 0; r < R; r++) fmc[r] = (c->rebuilt[r] && c->P.nstep_bp_mc > 0) ? 1 : fmc[r];
    }
    const int64_t st[3] = {c->istep, c->istep, stop};
    const int zero = 0;
    HIPCHK(hipMemcpyAsync(c->g_step_d.p, st, sizeof st, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->g_int_d.p, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(G.force_mc, fmc.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::fill(fmc.begin(), fmc.end(), 0);
    // buffers the captured kernels use are allocated before the capture
    const int nblk = force_blocks(c->N);
    c->nblk_force = nblk;
    if ((int64_t)c->blk_top2_d.n < (int64_t)R * nblk * 2) c->blk_top2_d.alloc((size_t)R * nblk * 2);
    if (c->P.nstep_bp_mc > 0 && c->seg_d.n < (size_t)R * c->N * SEGCAP) c->seg_d.alloc((size_t)R * c->N * SEGCAP);
    if (c->P.nstep_bp_mc > 0 && c->exc_list_d.n < (size_t)R * c->N + 1) c->exc_list_d.alloc((size_t)R * c->N + 1);
    // GRAPH_STEPS steps from the current position buffer, captured once
    // and replayed until something the launches take by value changes
    // (parameters, buffer addresses and capacities, the current buffer)
    PhaseGraphKey key;
    std::memset(&key, 0, sizeof key);
    key.P = c->P;
    key.D = c->ptrs();
    key.pos = c->pos_d.p;
    key.nblk = c->nblk_force;
    key.block = c->block;
    const bool perf_mc = c->P.nstep_bp_mc > 0;
    // one batch of steps: begin + resets, base pairs (+ MC), pair forces +
    // update, skin test; the host's buffer swaps return to the start
    auto launch_steps = [&]() {
      for (int j = 0; j < GRAPH_STEPS; j++) {
        hipLaunchKernelGGL(k_gstep_reset, grid_beads(c->N, R, 256), dim3(256), 0, c->stream, c->P, c->ptrs(), G,
                           c->exc_cnt_d.p, c->exc_fill_d.p, perf_mc ? c->exc_list_d.p : nullptr, perf_mc ? 1 : 0);
        bp_phase(c, nullptr, nullptr, G.active, G.istep_cur, true);
        if (c->P.nstep_bp_mc > 1 && c->D.maxbp > 0)
          hipLaunchKernelGGL(k_mc_keep, dim3((c->D.maxbp + 255) / 256), dim3(256), 0, c->stream, c->P, c->ptrs(),
                             -1, G.mc, G.active);
        force_kernel(c, 1, G.active, G.istep_cur);
        std::swap(c->pos_d, c->pos2_d);
        // a large system's thousands of block top-2 values over 1,024
        // threads (C4: 1.3% a step); a few hundred over 256
        if (c->nblk_force > 1024)
          hipLaunchKernelGGL(k_skin<1024>, dim3(R), dim3(1024), 0, c->stream, c->P, c->ptrs(), c->nblk_force, G);
        else
          hipLaunchKernelGGL(k_skin<256>, dim3(R), dim3(256), 0, c->stream, c->P, c->ptrs(), c->nblk_force, G);
      }
    };
    // issued directly: the host's launch cost overlaps the device's work,
    // which pays for large systems (a graph launch costs more per node
    // than a kernel launch here)
    if (c->pg_eager) {
      run_phase_batches(c, launch_steps);
      continue;
    }
    int slot = -1;
    for (int q = 0; q < 2; q++)
      if (c->pg_exec[q] && std::memcmp(&key, &c->pg_key[q], sizeof key) == 0) slot = q;
    if (slot >= 0) {
      hipGraphExec_t ex = c->pg_exec[slot];
      run_phase_batches(c, [&]() { HIPCHK(hipGraphLaunch(ex, c->stream)); });
      continue;
    }
    // the slot of the other buffer's graph is kept when it is still valid
    slot = 0;
    if (c->pg_exec[0]) {
      PhaseGraphKey k0 = c->pg_key[0];
      k0.pos = key.pos;
      k0.D.pos = key.D.pos;
      k0.D.pos_out = key.D.pos_out;
      if (std::memcmp(&k0, &key, sizeof key) == 0) slot = 1;
    }
    if (c->pg_exec[slot]) {
      HIPCHK(hipGraphExecDestroy(c->pg_exec[slot]));
      c->pg_exec[slot] = nullptr;
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    launch_steps();
    HIPCHK(hipStreamEndCapture(c->stream, &graph));
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(graph));
    c->pg_exec[slot] = exec;
    c->pg_key[slot] = key;
    run_phase_batches(c, [&]() { HIPCHK(hipGraphLaunch(exec, c->stream)); });
  }
}

}  // namespace

#include "sis_persistent_host.h"

// ======================================================================
// C ABI
// ======================================================================
#define API_BEGIN try {
#define API_END                                                      \
  }                                                                  \
  catch (const HipError &e) { return fail(SIS_ERR_HIP, e.msg); }     \
  catch (const ApiError &e) { return fail(e.code, e.msg); }          \
  catch (const std::exception &e) { return fail(SIS_ERR_ARG, e.what()); }

#include "sis_output_host.h"

// Device memory of sis_gpu_energy_noalloc: per-block partial sums, the
// output row and the call's conditions, mask and MC flags.  Each device
// the library runs on holds its own copy until the process ends; calls
// that overlap on one device share it and race.
static __device__ __align__(256) unsigned char scratch[SIS_NOALLOC_SCRATCH_BYTES];

namespace {

// Address of each device's `scratch` (the array's name in host code is a
// stand-in, not a device address), resolved once per device with that
// device current.
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

void resolve_scratch(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) throw ApiError{SIS_ERR_HIP, "device id beyond the scratch table"};
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw ApiError{SIS_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch array"};
  g_scratch_ptr[dev] = p;
}

void *scratch_of(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (dev >= 0 && dev < SCRATCH_MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
}

// conditions of one no-allocation energy call: temperature of every
// replica, the mask and MC flags selecting replica r0 (a kernel, not
// hipMemsetAsync, so that the call can sit in a captured graph)
__global__ void k_noalloc_setup(double *ecT, int *mask, int *mc, int R, int r0, double T) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  ecT[r] = T;
  mask[r] = r == r0 ? 1 : 0;
  mc[r] = 0;
}

}  // namespace

extern "C" {

const char *sis_gpu_last_error(void) { return g_err.c_str(); }

int sis_gpu_create(const sis_topo *topo, const sis_params *prm, const sis_options *opt,
                   sis_gpu_ctx **out) {
  API_BEGIN
  if (!topo || !prm || !out) return fail(SIS_ERR_ARG, "null argument");
  *out = nullptr;
  if (topo->nchains <= 0 || topo->nmp <= 0) return fail(SIS_ERR_ARG, "empty topology");
  // model 5 (triplet), or models 1/3 without the limit MC: their MC form
  // reads the base-pair type from the wrong list slot and model 4 never
  // forms a pair (src/force_bp_limit.F90:62, src/init_bp.F90:39)
  if (prm->bp_model != 5 && prm->bp_model != 0 && !((prm->bp_model == 1 || prm->bp_model == 3) && prm->nstep_bp_mc == 0))
    return fail(SIS_ERR_UNSUPPORTED, "base-pair model 5, or models 1 and 3 with nstep_bp_mc = 0");
  if (prm->max_bp_per_nt < 1) return fail(SIS_ERR_ARG, "max_bp_per_nt must be >= 1");
  if (prm->integrator < 0 || prm->integrator > 2)
    return fail(SIS_ERR_ARG, "integrator must be 1 (GJF-2GJ) or 2 (Ermak-McCammon)");
  sis_options o{};
  if (opt) o = *opt;
  if (o.nrep <= 0) o.nrep = 1;
  if (o.block != 0 && o.block != 128 && o.block != 256 && o.block != 512)
    return fail(SIS_ERR_ARG, "block: threads per replica of the batched form, 128, 256 or 512 (0 = 256)");
  if (o.grid < 0) return fail(SIS_ERR_ARG, "grid must be >= 0");
  if (o.flags & ~(SIS_OPT_HOST_EXCHANGE | SIS_OPT_NO_GRAPH)) return fail(SIS_ERR_ARG, "unknown option flags");
  int sum = 0;
  for (int c = 0; c < topo->nchains; c++) {
    if (topo->nmp_chain[c] < 1) return fail(SIS_ERR_ARG, "chain of zero length");
    sum += topo->nmp_chain[c];
  }
  if (sum != topo->nmp) return fail(SIS_ERR_ARG, "sum of chain lengths differs from nmp");
  int dev_count = 0;
  HIPCHK(hipGetDeviceCount(&dev_count));
  if (dev_count < 1) return fail(SIS_ERR_HIP, "no HIP device");
  auto *c = new sis_gpu_ctx();
  try {
    HIPCHK(hipGetDevice(&c->dev));
    resolve_scratch(c->dev);   // sis_gpu_energy_noalloc's scratch on this device
    c->opt = o;
    c->prm = *prm;
    c->N = topo->nmp;
    c->R = o.nrep;
    c->parity = o.parity;
    c->precision = o.precision;
    c->mode = o.mode;
    c->seed = o.seed ? o.seed : 0x5151u;
    const int N = c->N, R = c->R;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    // topology
    c->topo.resize(N);
    std::vector<uint8_t> charged(N, 1);
    for (int k = 0; k < topo->n_no_charge; k++) {
      int b = topo->no_charge[k] - 1;
      if (b < 0 || b >= N) throw ApiError{SIS_ERR_ARG, "no_charge index out of range"};
      charged[b] = 0;
    }
    int imp = 0;
    for (int ch = 0; ch < topo->nchains; ch++) {
      int n = topo->nmp_chain[ch];
      for (int i = 1; i <= n; i++, imp++) {
        int sq = topo->seq[imp];
        if (sq < 0 || sq > 4) throw ApiError{SIS_ERR_ARG, "sequence code out of range"};
        bool q = prm->dummy_has_charge ? true : sq != SQ_D;  // src/init_ele.F90:58-68
        if (!charged[imp]) q = false;
        c->topo[imp] = Topo{imp - i + 1, n, i, (topo->moltype[ch] == 2 ? 1 : 0) | (sq << 4) | ((q ? 1 : 0) << 8)};
      }
    }
    std::vector<int> cid(N);
    int maxlen = 1;
    for (int ch = 0, q = 0; ch < topo->nchains; ch++) {
      maxlen = std::max(maxlen, (int)topo->nmp_chain[ch]);
      for (int i = 0; i < topo->nmp_chain[ch]; i++) cid[q++] = ch;
    }
    {
      double span = (double)R * topo->nchains * (double)topo->nchains * maxlen * (double)maxlen;
      if (span > 9.0e18) throw ApiError{SIS_ERR_ARG, "system too large for the base-pair list key"};
    }
    c->chain_id_d.alloc(N);
    HIPCHK(hipMemcpy(c->chain_id_d.p, cid.data(), N * sizeof(int), hipMemcpyHostToDevice));
    c->P.nchains = topo->nchains;
    c->P.maxlen = maxlen;
    c->topo_d.alloc(N);
    HIPCHK(hipMemcpy(c->topo_d.p, c->topo.data(), N * sizeof(Topo), hipMemcpyHostToDevice));
    // parameters
    Params &P = c->P;
    P.bond_k = prm->bond_k; P.bond_r0 = prm->bond_r0;
    P.angl_k = prm->angl_k; P.angl_t0 = prm->angl_t0; P.angl_co0 = std::cos(prm->angl_t0);
    P.dih_k = prm->dih_k; P.dih_w = prm->dih_w; P.dih_p0 = prm->dih_p0;
    P.wca_sigma = prm->wca_sigma; P.wca_eps = prm->wca_eps;
    P.wca_sig2 = prm->wca_sigma * prm->wca_sigma;
    P.wca_coef = 12.0 * prm->wca_eps / P.wca_sig2;
    P.bp_cutoff_energy = prm->bp_cutoff_energy;
    P.nl_margin = prm->nl_margin;
    P.angl_reb = prm->angl_reb; P.dih_type = prm->dih_type; P.ele_on = prm->ele_on;
    P.ele_cut_type = prm->ele_cutoff_type; P.ele_cut_inp = prm->ele_cutoff;
    P.ele_excl_bond = prm->ele_excl_bond; P.ele_excl_angl = prm->ele_excl_angl;
    P.bp_model = prm->bp_model; P.bp_min_loop = prm->bp_min_loop;
    P.max_bp_per_nt = prm->max_bp_per_nt; P.nstep_bp_mc = prm->nstep_bp_mc; P.pbc = prm->pbc;
    P.N = N; P.R = R; P.fix_h4 = 0;
    P.bd = prm->integrator == 2;
    for (int t = 0; t < 3; t++) {
      const sis_bp_para &b = prm->bp[t];
      BpPara m;
      std::memcpy(&m, &b, sizeof m);
      std::memcpy(&c->bpp_h[t + 1], &b, sizeof m);
      BpPara mr = m;
      mr.angl_k1 = m.angl_k2; mr.angl_k2 = m.angl_k1; mr.angl_k3 = m.angl_k4; mr.angl_k4 = m.angl_k3;
      c->bpp_h[t + 4] = mr;
    }
    if (std::fabs(prm->bp_cutoff_energy) <= 2.220446049250313e-16) {
      P.bp_cutoff_dist = 18.0;
      for (int t = 1; t <= 6; t++) c->bpp_h[t].cutoff_ddist = P.bp_cutoff_dist - 13.8;
    } else {
      P.bp_cutoff_dist = 0.0;
      for (int t = 1; t <= 6; t++) {
        c->bpp_h[t].cutoff_ddist = std::sqrt(std::log(std::fabs(10.0 / prm->bp_cutoff_energy)) / c->bpp_h[t].bond_k);
        P.bp_cutoff_dist = std::max(P.bp_cutoff_dist, c->bpp_h[t].bond_r + c->bpp_h[t].cutoff_ddist);
      }
    }
    P.wca_nl_cut2 = (P.wca_sigma + P.nl_margin) * (P.wca_sigma + P.nl_margin);
    P.bp_nl_cut2 = (P.bp_cutoff_dist + P.nl_margin) * (P.bp_cutoff_dist + P.nl_margin);
    if (P.bp_model == 0) P.bp_nl_cut2 = -1.0;
    std::vector<int> hash;
    std::vector<double> dH, dS;
    build_triplet_table(*prm, hash, dH, dS);
    c->bp3_hash_d.alloc(hash.size());
    c->bp3_dH_d.alloc(dH.size());
    c->bp3_dS_d.alloc(dS.size());
    HIPCHK(hipMemcpy(c->bp3_hash_d.p, hash.data(), hash.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->bp3_dH_d.p, dH.data(), dH.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->bp3_dS_d.p, dS.data(), dS.size() * sizeof(double), hipMemcpyHostToDevice));
    // dynamics coefficients (src/job_md.F90:94-133, 700-767); the Brownian
    // integrator takes the viscosity in other units (lines 102-108) and
    // keeps only a = h/gamma (lines 725-731)
    c->mass.resize(N);
    c->fric.resize(N);
    double v = P.bd ? ((1.0 / 4184.0) * 6.02214076e23 * 1.0e-18) * prm->viscosity_pas
                    : prm->viscosity_pas * std::sqrt(1.0e3 / 4184.0) * 6.02214076e23 * 1.0e-20;
    std::vector<double> ma(N), mf(N), mx(N);
    for (int i = 0; i < N; i++) {
      c->mass[i] = md_mass(topo->seq[i]);
      c->fric[i] = 6.0 * PI_D * v * 10.0;
      if (P.bd) {
        ma[i] = prm->dt / c->fric[i];
        mf[i] = mx[i] = 0.0;
        continue;
      }
      double c1 = 0.5 * prm->dt * c->fric[i] / c->mass[i];
      double c2 = std::sqrt(1.0 / (1.0 + c1));
      ma[i] = (1.0 - c1) / (1.0 + c1);
      mf[i] = c2 * prm->dt / c->mass[i];
      mx[i] = c2 * prm->dt;
    }
    c->md_a_d.alloc(N); c->md_cf_d.alloc(N); c->md_cx_d.alloc(N); c->mass_d.alloc(N);
    HIPCHK(hipMemcpy(c->md_a_d.p, ma.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->md_cf_d.p, mf.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->md_cx_d.p, mx.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->mass_d.p, c->mass.data(), N * sizeof(double), hipMemcpyHostToDevice));
    // replicas (src/init_replica.F90:70-125, src/var_replica.F90:162-259)
    c->tK.assign(R, prm->tempK);
    c->ionic.assign(R, prm->ionic_strength);
    c->rep2lab.resize(R + 1); c->lab2rep.resize(R + 1); c->lab2val.resize(R + 1);
    for (int k = 1; k <= R; k++) { c->rep2lab[k] = k; c->lab2rep[k] = k; c->lab2val[k] = prm->tempK; }
    c->lab2ion.assign(R + 1, prm->ionic_strength);
    c->lab2rg.assign(R + 1, 0.0);
    c->lab2dcf.assign(R + 1, 0.0);
    c->pair_tb = make_pair_tb(R, std::vector<int>{R});
    c->pair_sets = 2;
    size_t RN = (size_t)R * N;
    c->md_cr_d.alloc(RN);
    c->tK_d.alloc(R); c->ele_coef_d.alloc(R); c->lambdaD_d.alloc(R); c->ele_cut_d.alloc(R);
    c->box_d.alloc(3 * R);
    c->bpp_d.alloc(7);
    HIPCHK(hipMemcpy(c->bpp_d.p, c->bpp_h, sizeof c->bpp_h, hipMemcpyHostToDevice));
    c->ecT_d.alloc(R); c->eCoef_d.alloc(R); c->eLam_d.alloc(R);
    c->pos_d.alloc(RN); c->pos2_d.alloc(RN); c->fbond_d.alloc(RN); c->vel_d.alloc(RN); c->acc_d.alloc(RN); c->move_d.alloc(RN); c->frc_d.alloc(RN);
    c->n_ele_d.alloc(RN); c->n_wca_d.alloc(RN); c->n_bp_tmp_d.alloc(RN); c->bp_off_d.alloc(RN + 1);
    c->n_bp_d.alloc(R); c->n_inc_d.alloc(RN); c->nt_cnt_d.alloc(RN); c->comp_d.alloc(RN);
    c->flags_d.alloc(8);
    c->blk_top2_d.alloc((size_t)R * ((N + 255) / 256 + 1) * 2);
    c->rebuild_flag_d.alloc(R); c->mask_d.alloc(R); c->mc_d.alloc(R); c->huge_d.alloc(R);
    c->scan_total_d.alloc(1);
    c->exc_cnt_d.alloc(RN); c->exc_off_d.alloc(RN + 1); c->exc_fill_d.alloc(RN); c->exc_list_d.alloc(RN + 1);   // exc_list: [0] count, then the components with excess
    c->ntlist_d.alloc((size_t)R * (N + 1));
    c->noise_d.alloc(RN * 3);
    c->bm_cache_d.alloc(2);
    c->ene_out_d.alloc((size_t)R * 16);
    c->tape_off_d.alloc(R + 1);
    c->tape_pos_d.alloc(R);
    c->tape_off_h.assign(R + 1, 0);
    // list capacities (grown on demand)
    c->D.maxnb = o.max_nbr > 0 ? o.max_nbr : 64;
    c->D.maxw = 16;
    c->D.maxbpb = 16;
    c->D.maxinc = 32;
    c->D.maxbp = std::max(64, N);
    alloc_lists(c);
    upload_conditions(c);
    setup_cells(c);
    *out = c;
  } catch (...) {
    delete c;
    throw;
  }
  return SIS_OK;
  API_END
}

int sis_gpu_destroy(sis_gpu_ctx *c) {
  API_BEGIN
  if (c) {
    dcd_shutdown(c);
    bp_shutdown(c);
    out_shutdown(c);
    rst_shutdown(c);
    twz_shutdown(c);
  }
  delete c;
  return SIS_OK;
  API_END
}

int sis_gpu_set_state(sis_gpu_ctx *c, const double *xyz, const double *velos, const double *accels) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  size_t RN = (size_t)c->R * c->N;
  std::vector<double4> h(RN);
  auto up = [&](const double *src, DevBuf<double4> &dst) {
    for (size_t g = 0; g < RN; g++) h[g] = double4{src[3 * g], src[3 * g + 1], src[3 * g + 2], 0.0};
    HIPCHK(hipMemcpy(dst.p, h.data(), RN * sizeof(double4), hipMemcpyHostToDevice));
  };
  if (xyz) { up(xyz, c->pos_d); c->lists_valid = false; }
  if (velos) up(velos, c->vel_d);
  if (accels) up(accels, c->acc_d);
  return SIS_OK;
  API_END
}

int sis_gpu_set_positions(sis_gpu_ctx *c, int32_t irep, const double *xyz) {
  API_BEGIN
  if (!c || !xyz) return fail(SIS_ERR_ARG, "null argument");
  if (irep < 1 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  const int N = c->N;
  std::vector<double4> h((size_t)N);
  for (int i = 0; i < N; i++) h[i] = double4{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0};
  HIPCHK(hipMemcpyAsync(c->pos_d.p + (size_t)(irep - 1) * N, h.data(), (size_t)N * sizeof(double4),
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SIS_OK;
  API_END
}

int sis_gpu_get_rng_tape_pos(sis_gpu_ctx *c, int32_t irep, int64_t *pos) {
  API_BEGIN
  if (!c || !pos) return fail(SIS_ERR_ARG, "null argument");
  if (!c->parity) return fail(SIS_ERR_STATE, "context was not created in parity mode");
  if (irep == -1) {
    *pos = c->rep_tape_pos;
    return SIS_OK;
  }
  if (irep < 1 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(pos, c->tape_pos_d.p + (irep - 1), sizeof(int64_t), hipMemcpyDeviceToHost));
  return SIS_OK;
  API_END
}

int sis_gpu_get_state(sis_gpu_ctx *c, double *xyz, double *velos, double *accels) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  size_t RN = (size_t)c->R * c->N;
  std::vector<double4> h(RN);
  auto down = [&](double *dst, DevBuf<double4> &src) {
    HIPCHK(hipMemcpy(h.data(), src.p, RN * sizeof(double4), hipMemcpyDeviceToHost));
    for (size_t g = 0; g < RN; g++) { dst[3 * g] = h[g].x; dst[3 * g + 1] = h[g].y; dst[3 * g + 2] = h[g].z; }
  };
  if (xyz) down(xyz, c->pos_d);
  if (velos) down(velos, c->vel_d);
  if (accels) down(accels, c->acc_d);
  return SIS_OK;
  API_END
}

int sis_gpu_set_rep_conditions(sis_gpu_ctx *c, const double *tempK, const double *ionic) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  // every value is checked before any state changes
  for (int r = 0; r < c->R; r++) {
    if (tempK && !(tempK[r] > 0.0)) return fail(SIS_ERR_ARG, "temperature must be positive");
    if (ionic && !(ionic[r] > 0.0)) return fail(SIS_ERR_ARG, "ionic strength must be positive");
  }
  if (tempK) {
    // new temperatures reset the labels to the identity: the values the
    // labels carry are re-keyed to the replicas that hold them now
    const int R = c->R;
    std::vector<double> ion(R), rg(R), dcf(R);
    for (int r = 0; r < R; r++) {
      const int lab = c->rep2lab[r + 1];
      ion[r] = c->lab2ion[lab]; rg[r] = c->lab2rg[lab]; dcf[r] = c->lab2dcf[lab];
    }
    for (int r = 0; r < R; r++) { c->lab2ion[r + 1] = ion[r]; c->lab2rg[r + 1] = rg[r]; c->lab2dcf[r + 1] = dcf[r]; }
  }
  for (int r = 0; r < c->R; r++) {
    if (tempK) {
      c->tK[r] = tempK[r];
      c->lab2val[r + 1] = tempK[r];
      c->rep2lab[r + 1] = r + 1;
      c->lab2rep[r + 1] = r + 1;
    }
    if (ionic) {
      c->ionic[r] = ionic[r];
      c->lab2ion[c->rep2lab[r + 1]] = ionic[r];   // the label carries it through exchanges
    }
  }
  upload_conditions(c);
  setup_cells(c);
  c->lists_valid = false;
  return SIS_OK;
  API_END
}

int sis_gpu_rebuild(sis_gpu_ctx *c, int32_t irep) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (irep < 0 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  std::vector<int> m(c->R, irep == 0 ? 1 : 0);
  if (irep > 0) m[irep - 1] = 1;
  if (!c->lists_valid || c->lists_analysis) std::fill(m.begin(), m.end(), 1);
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  rebuild(c, m);
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms[1] = ms;
  return SIS_OK;
  API_END
}

int sis_gpu_analysis_lists(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  analysis_rebuild(c);
  return SIS_OK;
  API_END
}

int sis_gpu_force(sis_gpu_ctx *c, int32_t irep, int32_t bp_mc, double *forces) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (irep < 1 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "force called before rebuild");
  std::vector<int> m(c->R, 0), mc(c->R, 0);
  m[irep - 1] = 1;
  mc[irep - 1] = bp_mc ? 1 : 0;
  set_mask(c, m);
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (c->mode == 3 && batched_ok(c) && !c->lists_analysis) {
    // a context created for the batched form evaluates with that kernel
    // (which builds the MD job's lists itself)
    batched_force(c, irep - 1, bp_mc);
  } else {
    bp_phase(c, &mc, nullptr, c->mask_d.p);
    force_kernel(c, 0, c->mask_d.p);
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  if (forces) {
    std::vector<double4> h(c->N);
    HIPCHK(hipMemcpyAsync(h.data(), c->frc_d.p + (size_t)(irep - 1) * c->N, c->N * sizeof(double4),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->N; i++) { forces[3 * i] = h[i].x; forces[3 * i + 1] = h[i].y; forces[3 * i + 2] = h[i].z; }
  }
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms[2] = ms;
  check_flags(c);
  return SIS_OK;
  API_END
}

int sis_gpu_energy(sis_gpu_ctx *c, int32_t irep, double tempK, int32_t flags, double *energies) {
  API_BEGIN
  if (!c || !energies) return fail(SIS_ERR_ARG, "null argument");
  if (irep < 1 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "energy called before rebuild");
  std::vector<int> m(c->R, 0);
  m[irep - 1] = 1;
  std::vector<double> T = c->tK;
  T[irep - 1] = tempK > 0 ? tempK : c->tK[irep - 1];
  std::vector<double> out((size_t)c->R * 16);
  energy_phase(c, m, T, flags, out.data());
  for (int k = 0; k < 11; k++) energies[k] = out[(size_t)(irep - 1) * 16 + k];
  return SIS_OK;
  API_END
}

int sis_gpu_energy_noalloc(sis_gpu_ctx *c, int32_t irep, double tempK, int32_t flags, double *energies) {
  API_BEGIN
  if (!c || !energies) return fail(SIS_ERR_ARG, "null argument");
  if (irep < 1 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  if (flags & ~3) return fail(SIS_ERR_ARG, "unknown flags");
  if (flags & 1) return fail(SIS_ERR_ARG, "the limit MC needs scratch that sis_gpu_energy_noalloc does not have; use sis_gpu_energy");
  const int N = c->N, R = c->R, B = 256;
  const int nblk = (N + B - 1) / B;
  // partial sums [nblk*8], output [16], temperatures [R] (doubles), then
  // mask [R] and MC flags [R] (ints)
  const size_t need = ((size_t)nblk * 8 + 16 + (size_t)R) * sizeof(double) + (size_t)2 * R * sizeof(int);
  if (need > (size_t)SIS_NOALLOC_SCRATCH_BYTES)
    return fail(SIS_ERR_CAPACITY, "sis_gpu_energy_noalloc needs " + std::to_string(need) +
                                      " bytes of scratch for this context; the maximum is " +
                                      std::to_string((size_t)SIS_NOALLOC_SCRATCH_BYTES));
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "energy called before rebuild");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != c->dev) return fail(SIS_ERR_STATE, "the context's device must be current");
  void *base = scratch_of(dev);
  if (!base) return fail(SIS_ERR_STATE, "scratch of this device not resolved");
  double *partial = static_cast<double *>(base);
  double *out = partial + (size_t)nblk * 8;
  double *ecT = out + 16;
  int *mask = reinterpret_cast<int *>(ecT + R);
  int *mc = mask + R;
  const int r0 = irep - 1;
  const double T = tempK > 0 ? tempK : c->tK[r0];
  double coef, lam, cut;
  host_set_ele(c->prm, c->anneal_on ? c->ele_T[r0] : T, c->ionic[r0], coef, lam, cut);
  hipLaunchKernelGGL(k_noalloc_setup, dim3((R + 255) / 256), dim3(256), 0, c->stream, ecT, mask, mc, R, r0, T);
  Ptrs D = c->ptrs();
  if (!(flags & 2)) {   // base-pair energies at T without the limit MC
    hipLaunchKernelGGL(k_reset_cnt, grid_beads(N, R, B), dim3(B), 0, c->stream, c->P, D, mask);
    if (c->nbp_total > 0) {
      if (c->precision == 1)
        hipLaunchKernelGGL(k_bp_eval<float>, dim3((c->nbp_total + B - 1) / B), dim3(B), 0, c->stream, c->P, D,
                           c->nbp_total, mc, ecT, mask, 0);
      else
        hipLaunchKernelGGL(k_bp_eval<double>, dim3((c->nbp_total + B - 1) / B), dim3(B), 0, c->stream, c->P, D,
                           c->nbp_total, mc, ecT, mask, 0);
    }
  }
  hipLaunchKernelGGL(k_energy<B>, dim3(nblk, 1), dim3(B), 0, c->stream, c->P, D, ecT, nullptr, nullptr, partial,
                     mask, 0, r0, coef, lam);
  const double *ext = nullptr;
  if (c->ext_on) {   // external terms into their own buffer (made by sis_gpu_set_external)
    if (c->X.rg) hipLaunchKernelGGL(k_rg_stats<256>, dim3(R), dim3(256), 0, c->stream, c->P, D, c->X);
    hipLaunchKernelGGL(k_external_energy<256>, dim3(R), dim3(256), 0, c->stream, c->P, D, c->X);
    ext = c->x_ene_d.p;
  }
  hipLaunchKernelGGL(k_energy_reduce, dim3(1), dim3(1), 0, c->stream, partial, nblk, out, ext, r0);
  HIPCHK(hipMemcpyAsync(energies, out, 11 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  check_flags(c);
  return SIS_OK;
  API_END
}

int sis_gpu_kinetic(sis_gpu_ctx *c, int32_t irep, double *ekin) {
  API_BEGIN
  if (!c || !ekin) return fail(SIS_ERR_ARG, "null argument");
  if (irep < 0 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  std::vector<double> e(c->R);
  kinetic_all(c, e.data());
  if (irep == 0) std::copy(e.begin(), e.end(), ekin);
  else *ekin = e[irep - 1];
  return SIS_OK;
  API_END
}

int sis_gpu_get_pair_count(sis_gpu_ctx *c, int32_t irep, int32_t kind, int64_t *n) {
  API_BEGIN
  if (!c || !n) return fail(SIS_ERR_ARG, "null argument");
  if (irep < 1 || irep > c->R || kind < 0 || kind > 2) return fail(SIS_ERR_ARG, "bad irep/kind");
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "lists not built");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int N = c->N;
  if (kind == 2) {
    int a = 0, b = 0;
    HIPCHK(hipMemcpy(&a, c->bp_off_d.p + (size_t)(irep - 1) * N, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&b, c->bp_off_d.p + (size_t)irep * N, sizeof(int), hipMemcpyDeviceToHost));
    *n = b - a;
    return SIS_OK;
  }
  int maxs = kind == 0 ? c->D.maxw : c->D.maxnb;
  std::vector<int> cnt(N), nb((size_t)maxs * N);
  HIPCHK(hipMemcpy(cnt.data(), (kind == 0 ? c->n_wca_d.p : c->n_ele_d.p) + (size_t)(irep - 1) * N,
                   N * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nb.data(), (kind == 0 ? c->nb_wca_d.p : c->nb_ele_d.p) + (size_t)(irep - 1) * maxs * N,
                   nb.size() * sizeof(int), hipMemcpyDeviceToHost));
  int64_t t = 0;
  for (int a = 0; a < N; a++)
    for (int k = 0; k < std::min(cnt[a], maxs); k++) t += nb[(size_t)k * N + a] > a;
  *n = t;
  return SIS_OK;
  API_END
}

int sis_gpu_get_pairs(sis_gpu_ctx *c, int32_t irep, int32_t kind, int32_t *out, int64_t cap, int64_t *n) {
  API_BEGIN
  if (!c || !out) return fail(SIS_ERR_ARG, "null argument");
  if (irep < 1 || irep > c->R || kind < 0 || kind > 2) return fail(SIS_ERR_ARG, "bad irep/kind");
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "lists not built");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int N = c->N;
  int64_t t = 0;
  if (kind == 2) {
    int a = 0, b = 0;
    HIPCHK(hipMemcpy(&a, c->bp_off_d.p + (size_t)(irep - 1) * N, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&b, c->bp_off_d.p + (size_t)irep * N, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> mp((size_t)(b - a) * 8 + 8);
    if (b > a) HIPCHK(hipMemcpy(mp.data(), c->bp_mp_d.p + (size_t)a * 8, (size_t)(b - a) * 8 * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < b - a; k++) {
      if (t < cap) for (int q = 0; q < 7; q++) out[7 * t + q] = mp[8 * (size_t)k + q];
      t++;
    }
    if (n) *n = t;
    if (t > cap) return fail(SIS_ERR_CAPACITY, "output capacity too small");
    return SIS_OK;
  }
  int maxs = kind == 0 ? c->D.maxw : c->D.maxnb;
  std::vector<int> cnt(N), nb((size_t)maxs * N);
  HIPCHK(hipMemcpy(cnt.data(), (kind == 0 ? c->n_wca_d.p : c->n_ele_d.p) + (size_t)(irep - 1) * N,
                   N * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nb.data(), (kind == 0 ? c->nb_wca_d.p : c->nb_ele_d.p) + (size_t)(irep - 1) * maxs * N,
                   nb.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (int a = 0; a < N; a++) {
    std::vector<int> js;
    for (int k = 0; k < std::min(cnt[a], maxs); k++) {
      int b = nb[(size_t)k * N + a];
      if (b > a) js.push_back(b);
    }
    std::sort(js.begin(), js.end());
    for (int b : js) {
      if (t < cap) { out[2 * t] = a + 1; out[2 * t + 1] = b + 1; }
      t++;
    }
  }
  if (n) *n = t;
  if (t > cap) return fail(SIS_ERR_CAPACITY, "output capacity too small");
  return SIS_OK;
  API_END
}

int sis_gpu_get_bp(sis_gpu_ctx *c, int32_t irep, int32_t *bp_mp, double *bp_coef, uint8_t *status,
                   double *ene, int64_t cap, int64_t *n) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (irep < 1 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "lists not built");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int N = c->N;
  int a = 0, b = 0;
  HIPCHK(hipMemcpy(&a, c->bp_off_d.p + (size_t)(irep - 1) * N, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&b, c->bp_off_d.p + (size_t)irep * N, sizeof(int), hipMemcpyDeviceToHost));
  int64_t m = b - a;
  if (n) *n = m;
  if (m > cap) return fail(SIS_ERR_CAPACITY, "output capacity too small");
  if (m == 0) return SIS_OK;
  std::vector<int> mp((size_t)m * 8);
  std::vector<double> cf((size_t)m * 2), en(m);
  std::vector<uint8_t> st(m);
  HIPCHK(hipMemcpy(mp.data(), c->bp_mp_d.p + (size_t)a * 8, mp.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(cf.data(), c->bp_coef_d.p + (size_t)a * 2, cf.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st.data(), c->bp_status_d.p + a, m, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(en.data(), c->ene_bp_d.p + a, m * sizeof(double), hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < m; k++) {
    if (bp_mp) for (int q = 0; q < 7; q++) bp_mp[7 * k + q] = mp[8 * k + q];
    if (bp_coef) { bp_coef[2 * k] = cf[2 * k]; bp_coef[2 * k + 1] = cf[2 * k + 1]; }
    if (status) status[k] = st[k];
    if (ene) ene[k] = en[k];
  }
  return SIS_OK;
  API_END
}

int sis_gpu_set_rng(sis_gpu_ctx *c, uint64_t seed) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  c->seed = seed;
  return SIS_OK;
  API_END
}

int sis_gpu_set_rng_tape(sis_gpu_ctx *c, int32_t irep, const uint32_t *words, int64_t n) {
  API_BEGIN
  if (!c || (!words && n > 0)) return fail(SIS_ERR_ARG, "null argument");
  if (!c->parity) return fail(SIS_ERR_STATE, "context was not created in parity mode");
  if (irep == -1) {
    c->rep_tape.assign(words, words + n);
    c->rep_tape_pos = 0;
    return SIS_OK;
  }
  if (irep < 1 || irep > c->R) return fail(SIS_ERR_ARG, "irep out of range");
  // rebuild the concatenated tape with this replica's words replaced
  const int R = c->R;
  std::vector<uint32_t> all;
  std::vector<int64_t> off(R + 1, 0);
  std::vector<uint32_t> old;
  if (c->tape_d.n) {
    old.resize(c->tape_d.n);
    HIPCHK(hipMemcpy(old.data(), c->tape_d.p, old.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  std::vector<int64_t> pos(R, 0);
  HIPCHK(hipMemcpy(pos.data(), c->tape_pos_d.p, R * sizeof(int64_t), hipMemcpyDeviceToHost));
  for (int r = 0; r < R; r++) {
    off[r] = (int64_t)all.size();
    if (r == irep - 1) {
      all.insert(all.end(), words, words + n);
      pos[r] = 0;
    } else if (!old.empty()) {
      all.insert(all.end(), old.begin() + c->tape_off_h[r], old.begin() + c->tape_off_h[r + 1]);
    }
  }
  off[R] = (int64_t)all.size();
  c->tape_d.alloc(std::max<size_t>(all.size(), 1));
  if (!all.empty()) HIPCHK(hipMemcpy(c->tape_d.p, all.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->tape_off_d.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->tape_pos_d.p, pos.data(), R * sizeof(int64_t), hipMemcpyHostToDevice));
  c->tape_off_h = off;
  return SIS_OK;
  API_END
}

int sis_gpu_init_velocities(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  const int N = c->N, R = c->R, B = 256;
  if (c->parity) {
    // per replica: 3N velocity normals then 3N acceleration normals
    // (src/job_md.F90:157-176); replicas in order share the Box-Muller cache
    for (int r = 0; r < R; r++) {
      std::vector<int> m(R, 0);
      m[r] = 1;
      set_mask(c, m);
      hipLaunchKernelGGL(k_noise_parity, dim3(1), dim3(1), 0, c->stream, c->P, c->ptrs(), 3 * N, 0, c->mask_d.p);
      hipLaunchKernelGGL(k_init_vel, dim3((N + B - 1) / B, R), dim3(B), 0, c->stream, c->P, c->ptrs(), 0, 1, c->seed,
                         c->mask_d.p);
      hipLaunchKernelGGL(k_noise_parity, dim3(1), dim3(1), 0, c->stream, c->P, c->ptrs(), 3 * N, 0, c->mask_d.p);
      hipLaunchKernelGGL(k_init_vel, dim3((N + B - 1) / B, R), dim3(B), 0, c->stream, c->P, c->ptrs(), 1, 1, c->seed,
                         c->mask_d.p);
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    check_flags(c);
  } else {
    hipLaunchKernelGGL(k_init_vel, dim3((N + B - 1) / B, R), dim3(B), 0, c->stream, c->P, c->ptrs(), 0, 0, c->seed,
                       nullptr);
    hipLaunchKernelGGL(k_init_vel, dim3((N + B - 1) / B, R), dim3(B), 0, c->stream, c->P, c->ptrs(), 1, 0, c->seed,
                       nullptr);
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return SIS_OK;
  API_END
}

int sis_gpu_step(sis_gpu_ctx *c, int64_t nsteps, int64_t *rebuilds) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (nsteps < 0) return fail(SIS_ERR_ARG, "nsteps < 0");
  int64_t r0 = c->rebuilds;
  c->tev_used = 0;
  if (c->lists_analysis) c->lists_valid = false;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  // each form rebuilds invalid lists itself and forces MC after it
  int mode = c->mode;
  if (mode == 0) mode = choose_mode(c);
  auto run = [&](int64_t n) {
    if (mode == 2 && !c->parity) {
      persistent_steps(c, n);
    } else if (mode == 3 && !c->parity) {
      batched_steps(c, n);
    } else if (graph_steps_ok(c)) {
      phases_graph_steps(c, n);
    } else {
      for (int64_t s = 0; s < n; s++) step_phases(c);
    }
  };
  // segments end at save steps and exchange steps; at such a step the
  // frames are queued first, then the replicas exchange, the order of the
  // save and exchange blocks of src/job_md.F90:485-570; restart files are
  // written last (src/job_md.F90:664-666)
  const int64_t ex = (c->opt.exchange_every > 0 && c->R > 1) ? c->opt.exchange_every : 0;
  const int64_t sv = (c->dcd && c->dcd->nstep_save > 0) ? c->dcd->nstep_save : 0;
  const int64_t sb = (c->bpo && c->bpo->nstep_save > 0) ? c->bpo->nstep_save : 0;
  const int64_t so = (c->outf && c->outf->nstep_save > 0) ? c->outf->nstep_save : 0;
  const int64_t rs = (c->rst && c->rst->nstep_save > 0) ? c->rst->nstep_save : 0;
  // whole segments of a batched replica-exchange run with nothing else at
  // their ends: exchange on the device, segments replayed as a HIP graph
  const bool dev_ex = mode == 3 && !c->parity && ex && device_exchange_ok(c) && !sv && !sb && !so && !rs &&
                      !c->rz_on && !c->twz;
  for (int64_t left = nsteps; left > 0;) {
    if (dev_ex && c->istep % ex == 0 && left >= ex) {
      left -= batched_device_run(c, left / ex);
      continue;
    }
    const int64_t rz = c->rz_on ? c->rz_step : 0;
    // .twz lines are taken before a step whose count is a multiple of the
    // cadence (the force of that step writes them)
    const int64_t tw = (c->twz && c->ext_on && c->X.nfr > 0) ? c->twz->nstep_save : 0;
    if (tw && (c->istep + 1) % tw == 0) twz_lines(c, c->istep + 1);
    // the next annealing entry is taken at the end of step an_next - 1
    // (src/job_md.F90:600-612)
    const bool an = c->anneal_on && c->an_next > 0 && c->an_next - 1 > c->istep;
    const bool rgs = c->ext_on && c->X.rg && c->rgs_next > 0 && c->rgs_next - 1 > c->istep;
    int64_t k = left;
    if (ex) k = std::min(k, ex - c->istep % ex);
    if (sv) k = std::min(k, sv - c->istep % sv);
    if (sb) k = std::min(k, sb - c->istep % sb);
    if (so) k = std::min(k, so - c->istep % so);
    if (rs) k = std::min(k, rs - c->istep % rs);
    if (an) k = std::min(k, c->an_next - 1 - c->istep);
    if (rz) k = std::min(k, rz - c->istep % rz);
    if (rgs) k = std::min(k, c->rgs_next - 1 - c->istep);
    if (tw) {   // up to the step before the next multiple
      const int64_t m = (c->istep + 1) % tw;
      k = std::min(k, m == 0 ? tw : tw - m);
    }
    run(k);
    left -= k;
    const bool exch = ex && c->istep % ex == 0;
    if (sv && c->istep % sv == 0) dcd_enqueue(c);
    if (sb && c->istep % sb == 0) bp_write_now(c);
    if (so && c->istep % so == 0) {
      // energy lines: labels, temperatures and Ekin before the exchange;
      // the energies of an exchange step are its own-condition energies
      // (fresh limit MC, src/job_md.F90:474-481), otherwise the step's
      // base-pair state is reused (flg_bp_energy)
      const int R = c->R;
      std::vector<int> lab(R);
      for (int r = 0; r < R; r++) lab[r] = c->rep2lab[r + 1];
      const std::vector<double> temps = c->tK;
      std::vector<double> ekin(R), e16((size_t)R * 16);
      kinetic_all(c, ekin.data());
      if (exch) {
        do_exchange(c, nullptr, nullptr, e16.data());
      } else {
        std::vector<int> all(R, 1);
        energy_phase(c, all, c->tK, (c->P.nstep_bp_mc > 0 ? 1 : 0) | 2, e16.data());
      }
      out_lines(c, e16.data(), ekin.data(), lab, temps);
    } else if (exch) {
      do_exchange(c, nullptr, nullptr);
    }
    if (rz && c->istep % rz == 0) box_resize_now(c);
    if (an && c->istep + 1 == c->an_next) {
      const int n = (int)c->an_T.size();
      c->ianneal++;
      std::fill(c->tK.begin(), c->tK.end(), c->an_T[c->ianneal - 1]);
      c->an_next = c->ianneal < n ? c->an_step[c->ianneal] : 0;
      upload_conditions(c);
    }
    if (rgs && c->istep + 1 == c->rgs_next) {   // src/job_md.F90:617-627
      c->irgs++;
      rg_schedule_apply(c);
    }
    if (rs && c->istep % rs == 0) rst_write_all(c, c->rst->paths);
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms[0] = ms;
  c->tcollect();
  check_flags(c);
  if (rebuilds) *rebuilds = c->rebuilds - r0;
  return SIS_OK;
  API_END
}

int sis_gpu_autotune(sis_gpu_ctx *c, const double *skins, int32_t nskin, int64_t nsteps, double *report) {
  API_BEGIN
  if (!c || !skins || nskin < 1 || nsteps < 1) return fail(SIS_ERR_ARG, "invalid argument");
  if (c->parity) return fail(SIS_ERR_UNSUPPORTED, "the autotuner runs the performance forms");
  autotune(c, skins, nskin, nsteps, report);
  return SIS_OK;
  API_END
}

int sis_gpu_set_rep_dims(sis_gpu_ctx *c, int32_t ndim, const int32_t *nvalues) {
  API_BEGIN
  if (!c || !nvalues) return fail(SIS_ERR_ARG, "null argument");
  if (ndim < 1 || ndim > 4) return fail(SIS_ERR_ARG, "ndim must be 1..4");
  int64_t prod = 1;
  std::vector<int> dims(ndim);
  for (int d = 0; d < ndim; d++) {
    if (nvalues[d] < 1) return fail(SIS_ERR_ARG, "a dimension needs at least one value");
    dims[d] = nvalues[d];
    prod *= nvalues[d];
  }
  if (prod != c->R) return fail(SIS_ERR_ARG, "the product of the dimensions differs from nrep");
  c->pair_tb = make_pair_tb(c->R, dims);
  c->pair_sets = 2 * ndim;
  c->idx_pair = 0;
  return SIS_OK;
  API_END
}

int sis_gpu_exchange(sis_gpu_ctx *c, double *replica_energies, int32_t *accepted) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "exchange called before rebuild");
  if (c->lists_analysis) rebuild(c, std::vector<int>(c->R, 1));
  do_exchange(c, replica_energies, accepted);
  return SIS_OK;
  API_END
}

int sis_gpu_set_step(sis_gpu_ctx *c, int64_t istep) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (istep < 0) return fail(SIS_ERR_ARG, "istep < 0");
  c->istep = istep;
  return SIS_OK;
  API_END
}

int sis_gpu_set_rep_dcf(sis_gpu_ctx *c, const double *k, const double *direction) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!k) { c->dcf_remd = false; return SIS_OK; }
  if (!direction) return fail(SIS_ERR_ARG, "null direction");
  if (!(c->ext_on && c->X.ndcf == 1)) return fail(SIS_ERR_STATE, "force exchange needs exactly one constant-force pair");
  const double n = std::sqrt(direction[0] * direction[0] + direction[1] * direction[1] + direction[2] * direction[2]);
  if (!(n > 0.0)) return fail(SIS_ERR_ARG, "zero force direction");
  for (int i = 0; i < 3; i++) c->dcf_dir[i] = direction[i] / n;   // src/init_tweezers.F90:32
  for (int r = 0; r < c->R; r++) c->lab2dcf[c->rep2lab[r + 1]] = k[r];
  c->dcf_remd = true;
  dcf_upload(c);
  return SIS_OK;
  API_END
}

int sis_gpu_set_rep_rg(sis_gpu_ctx *c, const double *rg0) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!rg0) { c->rg_remd = false; return SIS_OK; }
  if (!(c->ext_on && c->X.rg)) return fail(SIS_ERR_STATE, "the Rg bias is not set (sis_gpu_set_external)");
  const int R = c->R;
  for (int r = 0; r < R; r++) {
    if (!(rg0[r] > 0.0)) return fail(SIS_ERR_ARG, "Rg target must be positive");
    c->lab2rg[c->rep2lab[r + 1]] = rg0[r];
  }
  HIPCHK(hipMemcpyAsync(c->x_rg0_d.p, rg0, R * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->rg_remd = true;
  return SIS_OK;
  API_END
}

int sis_gpu_set_rg_schedule(sis_gpu_ctx *c, int32_t n, const int64_t *steps, const double *k, const double *r0) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (n <= 0) { c->rgs_next = 0; c->rgs_step.clear(); return SIS_OK; }
  if (!steps || !k || !r0) return fail(SIS_ERR_ARG, "null schedule");
  if (!(c->ext_on && c->X.rg)) return fail(SIS_ERR_STATE, "the Rg bias is not set (sis_gpu_set_external)");
  for (int i = 0; i < n; i++) {
    // read_timed_bias_rg's checks (src/read_timed_bias_rg.F90:31-37)
    if (steps[i] < 0 || k[i] < 0.0 || !(r0[i] > 0.0)) return fail(SIS_ERR_ARG, "invalid Rg-bias schedule entry");
    if (i > 0 && steps[i] <= steps[i - 1]) return fail(SIS_ERR_ARG, "Rg-bias schedule steps must increase");
  }
  c->rgs_step.assign(steps, steps + n);
  c->rgs_k.assign(k, k + n);
  c->rgs_0.assign(r0, r0 + n);
  // job_md's set-up (src/job_md.F90:231-265): the entry of the current step
  int i = 1;
  if (c->istep >= steps[n - 1]) i = n;
  else while (i < n && steps[i] <= c->istep) i++;
  c->irgs = i;
  rg_schedule_apply(c);
  return SIS_OK;
  API_END
}

int sis_gpu_set_box_resize(sis_gpu_ctx *c, int64_t resize_step, const double *change, const double *target) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (resize_step <= 0) { c->rz_on = false; return SIS_OK; }
  if (!c->prm.pbc) return fail(SIS_ERR_ARG, "box resizing needs a periodic box");
  if (!change) return fail(SIS_ERR_ARG, "null change");
  for (int k = 0; k < 3; k++) {
    c->rz_change[k] = change[k];
    c->rz_target[k] = target ? target[k] : 0.0;
    c->rz_orig[k] = c->prm.box[k];   // init_pbc: the box at the start (after a restart's PBC block)
    c->rz_count[k] = 0;
  }
  c->rz_has_target = target != nullptr;
  c->rz_step = resize_step;
  c->rz_on = true;
  return SIS_OK;
  API_END
}

int sis_gpu_set_anneal(sis_gpu_ctx *c, int32_t n, const int64_t *steps, const double *temps) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (n <= 0) {
    if (c->anneal_on) {
      c->anneal_on = false;
      c->P.mc_kT = 0.0;
      upload_conditions(c);
    }
    return SIS_OK;
  }
  if (!steps || !temps) return fail(SIS_ERR_ARG, "null schedule");
  for (int k = 0; k < n; k++) {
    if (!(temps[k] > 0.0)) return fail(SIS_ERR_ARG, "annealing temperature must be positive");
    if (k > 0 && steps[k] <= steps[k - 1]) return fail(SIS_ERR_ARG, "annealing steps must increase");
  }
  if (c->R > 1 && c->opt.exchange_every > 0)
    return fail(SIS_ERR_UNSUPPORTED, "annealing with temperature exchange");
  c->an_step.assign(steps, steps + n);
  c->an_T.assign(temps, temps + n);
  // job_md's set-up (src/job_md.F90:207-230): the entry of the current step
  // (from the restart file's ANNEAL block on a restarted run), the first
  // entry's temperature for the electrostatics, and the force's MC at the
  // input temperature's kT
  int ia = 1;
  if (c->istep >= c->an_step[n - 1]) ia = n;
  else if (c->rst_ianneal > 0) ia = std::min(c->rst_ianneal, n);
  c->ianneal = ia;
  c->an_next = ia < n ? c->an_step[ia] : 0;
  c->ele_T.assign(c->R, c->an_T[0]);
  std::fill(c->tK.begin(), c->tK.end(), c->an_T[ia - 1]);
  c->P.mc_kT = KB_KCAL * c->prm.tempK;
  c->anneal_on = true;
  upload_conditions(c);
  setup_cells(c);
  c->lists_valid = false;
  return SIS_OK;
  API_END
}

int sis_gpu_set_ss(sis_gpu_ctx *c, int32_t npair, const int32_t *pairs) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  const int N = c->N;
  if (npair < 0) {
    c->ss_on = false;
    c->lists_valid = false;
    return SIS_OK;
  }
  if (npair > 0 && !pairs) return fail(SIS_ERR_ARG, "null pairs");
  // the filters of read_ss (src/read_ss.F90:126-205) on the topology
  auto sq = [&](int a) { return topo_seq(c->topo[a]); };
  auto comp = [](int a, int b) {
    return (a == SQ_A && b == SQ_U) || (a == SQ_U && (b == SQ_A || b == SQ_G)) ||
           (a == SQ_G && (b == SQ_C || b == SQ_U)) || (a == SQ_C && b == SQ_G);
  };
  std::vector<std::vector<int>> part(N);
  for (int k = 0; k < npair; k++) {
    int imp = pairs[2 * k], jmp = pairs[2 * k + 1];
    if (jmp == 0) continue;
    if (imp < 1 || imp > N || jmp < 1 || jmp > N) return fail(SIS_ERR_ARG, "pair bead index out of range");
    if (jmp < imp) std::swap(imp, jmp);
    if (imp == jmp) continue;
    const int a = imp - 1, b = jmp - 1;
    const Topo &ta = c->topo[a], &tb = c->topo[b];
    if (ta.loc == 1 || ta.loc == ta.len || tb.loc == 1 || tb.loc == tb.len) continue;   // chain ends
    if (ta.start == tb.start && ta.loc + c->P.bp_min_loop >= tb.loc) continue;         // minimum loop
    if (c->P.bp_model >= 3 && !comp(sq(a - 1), sq(b + 1)) && !comp(sq(a + 1), sq(b - 1))) continue;   // isolated
    const int x = sq(a), y = sq(b);
    const bool canon = (x == SQ_G && (y == SQ_C || y == SQ_U)) || (x == SQ_C && y == SQ_G) ||
                       (x == SQ_A && y == SQ_U) || (x == SQ_U && (y == SQ_A || y == SQ_G));
    if (!canon) continue;
    part[a].push_back(b);
  }
  std::vector<int> off(N + 1, 0), flat;
  for (int a = 0; a < N; a++) {
    std::sort(part[a].begin(), part[a].end());
    part[a].erase(std::unique(part[a].begin(), part[a].end()), part[a].end());
    off[a] = (int)flat.size();
    flat.insert(flat.end(), part[a].begin(), part[a].end());
  }
  off[N] = (int)flat.size();
  c->ss_off_d.alloc(N + 1);
  HIPCHK(hipMemcpy(c->ss_off_d.p, off.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
  c->ss_partner_d.alloc(std::max<size_t>(flat.size(), 1));
  if (!flat.empty()) HIPCHK(hipMemcpy(c->ss_partner_d.p, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice));
  c->ss_on = true;
  c->lists_valid = false;
  return SIS_OK;
  API_END
}

int sis_gpu_set_external(sis_gpu_ctx *c, const sis_external *e) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  const int N = c->N, R = c->R;
  c->rgs_next = 0;   // a timed Rg bias is set again after the terms
  c->rgs_step.clear();
  if (!e || (!e->stage && e->n_dcf == 0 && e->n_fr == 0 && !e->rg && e->n_rest == 0)) {
    c->ext_on = false;
    c->X = ExtDev{};
    return SIS_OK;
  }
  if (e->n_dcf < 0 || e->n_fr < 0 || e->n_rest < 0 || e->rg < 0 || e->rg > 2)
    return fail(SIS_ERR_ARG, "invalid external-term counts");
  if (e->stage && !(e->stage_sigma > 0.0)) return fail(SIS_ERR_ARG, "stage_sigma must be positive");
  if ((e->n_dcf && (!e->dcf_pairs || !e->dcf_forces)) ||
      (e->n_fr && (!e->fr_pairs || !e->fr_k || !e->fr_init || !e->fr_velo)) || (e->rg && !e->rg_0) ||
      (e->n_rest && (!e->rest_id || !e->rest_rcut || !e->rest_para)))
    return fail(SIS_ERR_ARG, "null external-term array");
  auto bead = [&](int32_t b) {
    if (b < 1 || b > N) throw ApiError{SIS_ERR_ARG, "external-term bead index out of range"};
    return b - 1;
  };
  // per-bead lists of term ends, in the order the reference adds them
  // (tweezers: constant-force pairs, then ramps; then restraints)
  std::vector<std::vector<int2>> per(N);
  std::vector<int> dcfp(2 * e->n_dcf), frp(2 * e->n_fr), rid(3 * e->n_rest);
  for (int k = 0; k < e->n_dcf; k++) {
    dcfp[2 * k] = bead(e->dcf_pairs[2 * k]);
    dcfp[2 * k + 1] = bead(e->dcf_pairs[2 * k + 1]);
    per[dcfp[2 * k]].push_back(int2{XT_DCF1, k});
    per[dcfp[2 * k + 1]].push_back(int2{XT_DCF2, k});
  }
  for (int k = 0; k < e->n_fr; k++) {
    frp[2 * k] = bead(e->fr_pairs[2 * k]);
    frp[2 * k + 1] = bead(e->fr_pairs[2 * k + 1]);
    per[frp[2 * k]].push_back(int2{XT_FR1, k});
    per[frp[2 * k + 1]].push_back(int2{XT_FR2, k});
  }
  for (int k = 0; k < e->n_rest; k++) {
    const int ty = e->rest_id[3 * k + 2];
    if (ty != 1 && ty != 2) return fail(SIS_ERR_ARG, "restraint type must be 1 (sigmoid) or 2 (sigmoid-to-bead)");
    if (!(e->rest_para[3 * k + 2] != 0.0)) return fail(SIS_ERR_ARG, "restraint width must be non-zero");
    rid[3 * k] = bead(e->rest_id[3 * k]);
    rid[3 * k + 1] = bead(e->rest_id[3 * k + 1]);
    rid[3 * k + 2] = ty;
    per[rid[3 * k]].push_back(int2{XT_REST_I, k});
    if (ty == 1) per[rid[3 * k + 1]].push_back(int2{XT_REST_J, k});
  }
  std::vector<int> off(N + 1, 0);
  std::vector<int2> terms;
  for (int a = 0; a < N; a++) {
    off[a] = (int)terms.size();
    terms.insert(terms.end(), per[a].begin(), per[a].end());
  }
  off[N] = (int)terms.size();
  auto up = [&](auto &buf, const auto *src, size_t n) {
    buf.alloc(std::max<size_t>(n, 1));
    if (n) HIPCHK(hipMemcpy(buf.p, src, n * sizeof(*src), hipMemcpyHostToDevice));
    return buf.p;
  };
  ExtDev X{};
  X.stage = e->stage != 0; X.rg = e->rg; X.ndcf = e->n_dcf; X.nfr = e->n_fr; X.nrest = e->n_rest;
  X.stage_sigma = e->stage_sigma; X.stage_eps = e->stage_eps; X.rg_k = e->rg_k;
  std::vector<double> rg0(R, 0.0);
  if (e->rg) std::copy(e->rg_0, e->rg_0 + R, rg0.begin());
  c->rg_remd = false;   // ladders over the Rg target or the force are set again after the terms
  c->dcf_remd = false;
  X.rg0 = up(c->x_rg0_d, rg0.data(), (size_t)R);
  X.term_off = up(c->x_term_off_d, off.data(), off.size());
  X.terms = up(c->x_terms_d, terms.data(), terms.size());
  X.dcf_pairs = up(c->x_dcf_pairs_d, dcfp.data(), dcfp.size());
  X.dcf_f = up(c->x_dcf_f_d, e->dcf_forces, (size_t)3 * e->n_dcf * R);
  X.fr_pairs = up(c->x_fr_pairs_d, frp.data(), frp.size());
  X.fr_k = up(c->x_fr_k_d, e->fr_k, (size_t)2 * e->n_fr);
  X.fr_init = up(c->x_fr_init_d, e->fr_init, (size_t)6 * e->n_fr);
  X.fr_velo = up(c->x_fr_velo_d, e->fr_velo, (size_t)6 * e->n_fr);
  X.rest_id = up(c->x_rest_id_d, rid.data(), rid.size());
  X.rest_rcut = up(c->x_rest_rcut_d, e->rest_rcut, (size_t)e->n_rest);
  X.rest_para = up(c->x_rest_para_d, e->rest_para, (size_t)3 * e->n_rest);
  c->x_rg_stat_d.alloc((size_t)4 * R);
  c->x_ene_d.alloc((size_t)4 * R);
  c->x_err_d.alloc(1);
  X.rg_stat = c->x_rg_stat_d.p;
  X.ene = c->x_ene_d.p;
  X.err = c->x_err_d.p;
  c->X = X;
  c->ext_on = true;
  return SIS_OK;
  API_END
}

int sis_gpu_set_labels(sis_gpu_ctx *c, const int32_t *rep2lab) {
  API_BEGIN
  if (!c || !rep2lab) return fail(SIS_ERR_ARG, "null argument");
  const int R = c->R;
  std::vector<int> seen(R + 1, 0);
  for (int r = 0; r < R; r++) {
    if (rep2lab[r] < 1 || rep2lab[r] > R || seen[rep2lab[r]]) return fail(SIS_ERR_ARG, "labels are not a permutation of 1..nrep");
    seen[rep2lab[r]] = 1;
  }
  for (int r = 0; r < R; r++) {
    c->rep2lab[r + 1] = rep2lab[r];
    c->lab2rep[rep2lab[r]] = r + 1;
    c->tK[r] = c->lab2val[rep2lab[r]];
    c->ionic[r] = c->lab2ion[rep2lab[r]];
  }
  if (c->ext_on && c->X.rg && c->rg_remd) {
    std::vector<double> g(R);
    for (int r = 0; r < R; r++) g[r] = c->lab2rg[rep2lab[r]];
    HIPCHK(hipMemcpyAsync(c->x_rg0_d.p, g.data(), R * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (c->ext_on && c->dcf_remd) dcf_upload(c);
  // the pair table restarts from its first half, as var_replica's idx_pair
  // does on a restarted run (it is not in the restart file,
  // src/var_replica.F90:85, src/write_rst.F90:40-52)
  c->idx_pair = 0;
  upload_conditions(c);
  setup_cells(c);
  c->lists_valid = false;
  return SIS_OK;
  API_END
}

int sis_gpu_get_labels(sis_gpu_ctx *c, int32_t *rep2lab) {
  API_BEGIN
  if (!c || !rep2lab) return fail(SIS_ERR_ARG, "null argument");
  for (int r = 0; r < c->R; r++) rep2lab[r] = c->rep2lab[r + 1];
  return SIS_OK;
  API_END
}

int sis_gpu_get_info(sis_gpu_ctx *c, int64_t *out, int32_t n) {
  API_BEGIN
  if (!c || !out) return fail(SIS_ERR_ARG, "null argument");
  int mode_used = c->parity ? 1 : (c->mode ? c->mode : choose_mode(c));
  if ((mode_used == 3 && !batched_ok(c)) || (mode_used == 2 && (c->R != 1 || c->ext_on))) mode_used = 1;   // forms fall back
  int64_t v[11] = {c->N, c->R, c->istep, c->rebuilds, 0, 0, 0, c->D.maxnb, mode_used, c->block, c->dev_exchanges};
  if (c->lists_valid) {
    int64_t k;
    if (sis_gpu_get_pair_count(c, 1, 0, &k) == SIS_OK) v[4] = k;
    if (sis_gpu_get_pair_count(c, 1, 1, &k) == SIS_OK) v[5] = k;
    if (sis_gpu_get_pair_count(c, 1, 2, &k) == SIS_OK) v[6] = k;
  }
  for (int k = 0; k < n && k < 11; k++) out[k] = v[k];
  return SIS_OK;
  API_END
}

double sis_gpu_last_kernel_ms(sis_gpu_ctx *c, int32_t which) {
  if (!c || which < 0 || which > 4) return -1.0;
  if (which == 3) return c->dom_launches ? c->dom_ms / (double)c->dom_launches : 0.0;
  if (which == 4) return (double)c->dom_launches;
  return c->last_ms[which];
}

}  // extern "C"

#include "sis_pool_host.h"
