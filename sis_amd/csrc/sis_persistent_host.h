// Host drivers of the persistent (mode 2) and batched (mode 3) step forms.
// Included by sis_gpu.hip after the per-phase driver.
#pragma once


bool batched_ok(sis_gpu_ctx *c) {
  return !c->parity && !c->ext_on && c->R >= 2 && c->P.nchains == 1 && c->N <= 1024;
}

// The form a context runs when none is asked for: the batched form for
// many replicas of a small chain, else the per-phase kernels (replayed
// from captured graphs in the performance form).  The persistent form
// stays available on request; on one large system the replayed per-phase
// kernels are faster (each at its own occupancy, no grid barriers).
int choose_mode(sis_gpu_ctx *c) {
  if (c->parity) return 1;
  if (batched_ok(c)) return 3;
  return 1;
}

namespace {

int persistent_grid(sis_gpu_ctx *c, const void *kfn, int B) {
  int dev = 0, ncu = 0, per = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, B, 0));
  if (per < 1) throw ApiError{SIS_ERR_HIP, "persistent kernel does not fit on a compute unit"};
  int want = (c->N * 8 + B - 1) / B;               // one 8-lane group per bead
  int cap = per * ncu;                               // co-resident limit
  int g = std::min(want, cap);
  if (c->opt.grid > 0) g = std::min(cap, c->opt.grid);
  return std::max(g, 1);
}

}  // namespace

void persistent_steps(sis_gpu_ctx *c, int64_t nsteps) {
  if (c->R != 1 || c->ext_on) {   // external terms run in the per-phase form
    for (int64_t s = 0; s < nsteps; s++) step_phases(c);
    return;
  }
  const int N = c->N;
  constexpr int B = 256, LPB = 8;
  const void *kfn = c->precision == 1 ? (const void *)k_persistent<float, LPB, B>
                                      : (const void *)k_persistent<double, LPB, B>;
  const int G = persistent_grid(c, kfn, B);
  c->block = G;
  if ((int64_t)c->seg_d.n < (int64_t)N * SEGCAP) c->seg_d.alloc((size_t)N * SEGCAP);
  if ((int64_t)c->blk_top2_d.n < 2 * G) c->blk_top2_d.alloc(2 * (size_t)G);
  if (c->pers_sync_d.n < 4) c->pers_sync_d.alloc(4);
  if (c->pers_bar_d.n < 1) c->pers_bar_d.alloc(1);
  int pending = 0;
  HIPCHK(hipMemcpyAsync(&pending, c->rebuild_flag_d.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int64_t left = nsteps;
  int force_mc = 0;
  std::vector<int> all(1, 1);
  while (left > 0) {
    if (pending || !c->lists_valid) {
      rebuild(c, all);
      c->rebuilds++;
      force_mc = c->P.nstep_bp_mc > 0 ? 1 : 0;
      pending = 0;
    }
    Ptrs D = c->ptrs();
    hipLaunchKernelGGL(k_comp_init, grid_beads(N, 1, 256), dim3(256), 0, c->stream, c->P, D, c->exc_cnt_d.p,
                       c->exc_fill_d.p, nullptr);
    hipLaunchKernelGGL(k_reset_cnt, grid_beads(N, 1, 256), dim3(256), 0, c->stream, c->P, D, nullptr);
    HIPCHK(hipMemsetAsync(c->pers_bar_d.p, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(hipMemsetAsync(c->flags_d.p, 0, 8 * sizeof(int), c->stream));
    PersArgs A{};
    A.nsteps = left;
    A.step0 = c->istep;
    A.seed = c->seed;
    A.bar = c->pers_bar_d.p;
    A.seg_n = c->exc_cnt_d.p;
    A.seg = c->seg_d.p;
    A.has_exc = c->exc_fill_d.p;
    A.done = c->pers_sync_d.p;
    A.pos_b = c->pos2_d.p;
    A.nbp_total = c->nbp_total;
    A.force_mc_first = force_mc;
    A.huge = c->huge_d.p;
    Params P = c->P;
    void *args[] = {&P, &D, &A};
    c->tmark();
    HIPCHK(hipLaunchCooperativeKernel(kfn, dim3(G), dim3(B), args, 0, c->stream));
    c->tmark();
    int done[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(done, c->pers_sync_d.p, sizeof done, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    check_flags(c);
    if (done[0] <= 0) throw ApiError{SIS_ERR_STATE, "persistent kernel made no progress"};
    c->istep += done[0];
    left -= done[0];
    if (done[2]) std::swap(c->pos_d, c->pos2_d);
    pending = done[1];
    force_mc = 0;
  }
  HIPCHK(hipMemcpyAsync(c->rebuild_flag_d.p, &pending, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

namespace {

BpArrays packed_arrays(sis_gpu_ctx *c) {
  return BpArrays{c->bp_mp_d.p, c->bp_coef_d.p, c->bp_status_d.p, c->bp_status_mc_d.p, c->ene_bp_d.p,
                  c->for_bp_d.p};
}
BpArrays batched_arrays(sis_gpu_ctx *c) {
  return BpArrays{c->bb_mp_d.p, c->bb_coef_d.p, c->bb_status_d.p, c->bb_status_mc_d.p, c->bb_ene_d.p,
                  c->bb_for_d.p};
}

// candidate range of each replica in the packed layout
std::vector<int> packed_rep_off(sis_gpu_ctx *c) {
  const int N = c->N, R = c->R;
  std::vector<int> off((size_t)R * N + 1), ro(R + 1);
  HIPCHK(hipMemcpyAsync(off.data(), c->bp_off_d.p, off.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int r = 0; r <= R; r++) ro[r] = off[(size_t)r * N];
  return ro;
}

// batched slots per replica for the largest replica list `mx`
void ensure_batched_capacity(sis_gpu_ctx *c, int mx) {
  const int R = c->R;
  int capb = std::max(c->batch_capb, std::max(256, 2 * mx + 64));
  if ((int64_t)R * capb * 8 >= (int64_t)INT32_MAX)
    throw ApiError{SIS_ERR_CAPACITY, "batched base-pair candidate ids exceed 32 bits"};
  if (capb != c->batch_capb || c->bb_mp_d.n < (size_t)R * capb * 8) {
    c->batch_capb = capb;
    size_t RC = (size_t)R * capb;
    c->bb_mp_d.alloc(RC * 8);
    c->bb_coef_d.alloc(RC * 2);
    c->bb_status_d.alloc(RC);
    c->bb_status_mc_d.alloc(RC);
    c->bb_ene_d.alloc(RC);
    c->bb_for_d.alloc(RC * 18);
  }
}

// packed -> batched: candidates, incidences, pending rebuild flags
void to_batched(sis_gpu_ctx *c) {
  const int N = c->N, R = c->R;
  std::vector<int> ro = packed_rep_off(c), cnt(R);
  int mx = 0;
  for (int r = 0; r < R; r++) { cnt[r] = ro[r + 1] - ro[r]; mx = std::max(mx, cnt[r]); }
  ensure_batched_capacity(c, mx);
  const int capb = c->batch_capb;
  HIPCHK(hipMemcpyAsync(c->bb_off_d.p, ro.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->bb_nbp_d.p, cnt.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
  if (mx > 0)
    hipLaunchKernelGGL(k_bp_relayout, dim3((mx + 255) / 256, R), dim3(256), 0, c->stream, packed_arrays(c),
                       batched_arrays(c), c->bb_off_d.p, capb, 1);
  hipLaunchKernelGGL(k_inc_remap, grid_beads(N, R, 256), dim3(256), 0, c->stream, c->P, c->ptrs(), c->bb_off_d.p,
                     capb, 1);
  HIPCHK(hipMemcpyAsync(c->bb_rebuild_d.p, c->rebuild_flag_d.p, R * sizeof(int), hipMemcpyDeviceToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

// batched -> packed (the layout every other entry point reads)
void to_packed(sis_gpu_ctx *c) {
  const int N = c->N, R = c->R;
  const int capb = c->batch_capb;
  std::vector<int> cnt(R), ro(R + 1, 0);
  HIPCHK(hipMemcpyAsync(cnt.data(), c->bb_nbp_d.p, R * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int mx = 0;
  for (int r = 0; r < R; r++) { ro[r + 1] = ro[r] + cnt[r]; mx = std::max(mx, cnt[r]); }
  const int total = ro[R];
  if (total > c->D.maxbp) {
    c->D.maxbp = total + total / 4 + 64;
    alloc_bp(c);
  }
  HIPCHK(hipMemcpyAsync(c->bb_off_d.p, ro.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
  if (mx > 0)
    hipLaunchKernelGGL(k_bp_relayout, dim3((mx + 255) / 256, R), dim3(256), 0, c->stream, batched_arrays(c),
                       packed_arrays(c), c->bb_off_d.p, capb, 0);
  hipLaunchKernelGGL(k_inc_remap, grid_beads(N, R, 256), dim3(256), 0, c->stream, c->P, c->ptrs(), c->bb_off_d.p,
                     capb, 0);
  // per-bead candidate offsets of the packed layout
  scan_exclusive(c, c->n_bp_tmp_d.p, c->bp_off_d.p, (int64_t)R * N, nullptr);
  HIPCHK(hipMemcpyAsync(c->rebuild_flag_d.p, c->bb_rebuild_d.p, R * sizeof(int), hipMemcpyDeviceToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->nbp_total = total;
  c->bp_rep_off.assign(ro.begin(), ro.end());
  c->lists_valid = true;
}

// host rebuild of the replicas in `mask`; clears their pending flags and
// returns which replicas the first step must run the limit MC on
std::vector<int> host_rebuild(sis_gpu_ctx *c, const std::vector<int> &mask) {
  const int R = c->R;
  rebuild(c, mask);
  c->rebuilds++;
  std::vector<int> pend(R);
  HIPCHK(hipMemcpy(pend.data(), c->rebuild_flag_d.p, R * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> mcf(R, 0);
  for (int r = 0; r < R; r++)
    if (c->rebuilt[r]) { pend[r] = 0; mcf[r] = c->P.nstep_bp_mc > 0 ? 1 : 0; }
  HIPCHK(hipMemcpy(c->rebuild_flag_d.p, pend.data(), R * sizeof(int), hipMemcpyHostToDevice));
  return mcf;
}

}  // namespace

namespace {

void ensure_batched_io(sis_gpu_ctx *c) {
  const int R = c->R;
  const size_t RN = (size_t)R * c->N;
  if (c->seg_d.n < RN * SEGCAP) c->seg_d.alloc(RN * SEGCAP);
  if (c->bb_done_d.n < (size_t)R) {
    c->bb_done_d.alloc(R); c->bb_nbp_d.alloc(R); c->bb_rebuild_d.alloc(R); c->bb_todo_d.alloc(R);
    c->bb_stepno_d.alloc(R); c->bb_off_d.alloc(R + 1); c->bb_mcfirst_d.alloc(R); c->bb_nreb_d.alloc(R);
  }
}

// threads per replica's workgroup: 256 unless the options ask for 128 or 512
int batched_block(sis_gpu_ctx *c) { return c->opt.block == 128 || c->opt.block == 512 ? c->opt.block : 256; }

const void *batched_kernel(sis_gpu_ctx *c, int B) {
  const bool f = c->precision == 1;
  if (B == 512) return f ? (const void *)k_batched<float, 512> : (const void *)k_batched<double, 512>;
  if (B == 128) return f ? (const void *)k_batched<float, 128> : (const void *)k_batched<double, 128>;
  return f ? (const void *)k_batched<float, 256> : (const void *)k_batched<double, 256>;
}

// Dynamic LDS a k_batched launch may take beyond its static arrays: the
// device's per-workgroup maximum (160 KB on gfx950) less the kernel's
// static LDS.  The kernel is told it may use that much.
size_t batched_lds_cap(sis_gpu_ctx *c, int B) {
  const void *fn = batched_kernel(c, B);
  int dev = 0, maxb = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  hipFuncAttributes fa{};
  HIPCHK(hipFuncGetAttributes(&fa, fn));
  const size_t cap = (size_t)maxb > fa.sharedSizeBytes ? (size_t)maxb - fa.sharedSizeBytes : 0;
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap);
  return cap;
}

// arguments of one k_batched launch over the batched layout
BatchArgs batched_args(sis_gpu_ctx *c, int lpb) {
  const int N = c->N;
  BatchArgs A{};
  A.todo = c->bb_todo_d.p;
  A.stepno = c->bb_stepno_d.p;
  A.seed = c->seed;
  A.seg_n = c->exc_cnt_d.p;
  A.seg = c->seg_d.p;
  A.has_exc = c->exc_fill_d.p;
  A.done = c->bb_done_d.p;
  A.nbp = c->bb_nbp_d.p;
  A.rebuild = c->bb_rebuild_d.p;
  A.mc_first = c->bb_mcfirst_d.p;
  A.nreb = c->bb_nreb_d.p;
  A.pos_b = c->pos2_d.p;
  A.capb = c->batch_capb;
  A.lpb = lpb;
  A.xlds = N <= BATCH_X_LDS ? N : 0;
  A.fplds = (A.xlds && N * lpb <= BATCH_FP_LDS) ? N * lpb : 0;
  // the dynamics state for the launch, then the WCA and Debye-Hueckel
  // lists (at most their slot capacity), when they fit beside the positions
  // and partial forces
  A.nblds = 0;
  A.stlds = 0;
  A.mplds = 0;
  if (A.xlds) {
    const size_t cap = batched_lds_cap(c, batched_block(c));
    if (batched_lds_bytes(A.xlds, A.fplds, 0, N, 1) <= cap) A.stlds = 1;
    // the candidate records (up to the slot capacity)
    const size_t mpw = (size_t)c->batch_capb * 8;
    if (batched_lds_bytes(A.xlds, A.fplds, 0, N, A.stlds, (int)mpw) + (size_t)2 * N * sizeof(int) <= cap)
      A.mplds = (int)mpw;
    const size_t used = batched_lds_bytes(A.xlds, A.fplds, 0, N, A.stlds, A.mplds) + (size_t)2 * N * sizeof(int);
    const size_t want = (size_t)N * (c->D.maxw + (c->P.ele_on ? c->D.maxnb : 0));
    if (cap > used) A.nblds = (int)std::min(want, (cap - used) / sizeof(int));
  }
  A.force_only = 0;
  return A;
}
Ptrs batched_ptrs(sis_gpu_ctx *c) {
  Ptrs D = c->ptrs();
  D.bp_mp = c->bb_mp_d.p;
  D.bp_coef = c->bb_coef_d.p;
  D.bp_status = c->bb_status_d.p;
  D.bp_status_mc = c->bb_status_mc_d.p;
  D.ene_bp = c->bb_ene_d.p;
  D.for_bp = c->bb_for_d.p;
  return D;
}

}  // namespace

// The forces of replica irep0's current positions from the batched kernel
// itself (k_batched in force-only mode: its base-pair evaluation, limit MC
// when bp_mc, pair and bonded arithmetic in the context's precision), into
// frc_d.  Lists must be valid; nothing moves.
void batched_force(sis_gpu_ctx *c, int irep0, int bp_mc) {
  const int R = c->R;
  const int B = batched_block(c);
  ensure_batched_io(c);
  to_batched(c);
  std::vector<int> todo(R, 0), mcf(R, 0);
  std::vector<int64_t> stepno(R, c->istep);
  todo[irep0] = 1;
  mcf[irep0] = bp_mc ? 1 : 0;
  HIPCHK(hipMemcpyAsync(c->bb_todo_d.p, todo.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->bb_stepno_d.p, stepno.data(), R * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->bb_mcfirst_d.p, mcf.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->flags_d.p, 0, 8 * sizeof(int), c->stream));
  BatchArgs A = batched_args(c, lanes_per_bead(c->N));
  A.force_only = 1;
  Params P = c->P;
  Ptrs D = batched_ptrs(c);
  void *args[] = {&P, &D, &A};
  HIPCHK(hipLaunchKernel(batched_kernel(c, B), dim3(R), dim3(B), args,
                         batched_lds_bytes(A.xlds, A.fplds, A.nblds, c->N, A.stlds, A.mplds), c->stream));
  to_packed(c);
}

// Batched form: one workgroup per replica runs whole segments of steps
// (up to the next exchange).  Lists move to the batched layout for the
// segment and back afterwards, so skin counters, pending rebuilds and the
// limit-MC state carry over exactly as in the per-phase form.
namespace {

// Runs the batched kernel over the batched layout until every replica has
// done its `todo` steps (from `stepno`), growing lists through host
// rebuilds when an in-kernel rebuild outgrew a capacity.  mcf: replicas
// whose first step does the limit MC.
void batched_segment(sis_gpu_ctx *c, int64_t seg, std::vector<int> todo, std::vector<int64_t> stepno,
                     std::vector<int> mcf) {
  const int N = c->N, R = c->R;
  const int B = batched_block(c);
  const int lpb = lanes_per_bead(N);
  const void *kfn = batched_kernel(c, B);
  std::vector<int> done(R);
  for (int r = 0; r < R; r++) done[r] = (int)(seg - todo[r]);
  for (int attempt = 0;; attempt++) {
    HIPCHK(hipMemcpyAsync(c->bb_todo_d.p, todo.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->bb_stepno_d.p, stepno.data(), R * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->bb_mcfirst_d.p, mcf.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->flags_d.p, 0, 8 * sizeof(int), c->stream));
    BatchArgs A = batched_args(c, lpb);
    Params P = c->P;
    Ptrs D = batched_ptrs(c);
    void *args[] = {&P, &D, &A};
    c->tmark();
    HIPCHK(hipLaunchKernel(kfn, dim3(R), dim3(B), args, batched_lds_bytes(A.xlds, A.fplds, A.nblds, N, A.stlds, A.mplds),
                           c->stream));
    c->tmark();
    std::vector<int> d(R), nr(R), f(8);
    HIPCHK(hipMemcpyAsync(d.data(), c->bb_done_d.p, R * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(nr.data(), c->bb_nreb_d.p, R * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(f.data(), c->flags_d.p, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (f[7]) throw ApiError{SIS_ERR_STATE, "limit MC met a base-pair candidate outside the system"};
    if (f[1] == 2) throw ApiError{SIS_ERR_STATE, "limit MC found an excess nucleotide without pairs"};
    if (f[5]) throw ApiError{SIS_ERR_CAPACITY, "a base-pair conflict component outgrew the limit-MC scratch"};
    bool lag = false;
    std::fill(mcf.begin(), mcf.end(), 0);
    std::vector<int> lagm(R, 0);
    for (int r = 0; r < R; r++) {
      done[r] += d[r];
      stepno[r] += d[r];
      todo[r] = (int)(seg - done[r]);
      c->rebuilds += nr[r];
      if (todo[r] > 0) { lag = true; lagm[r] = 1; }
    }
    if (!lag) break;
    if (attempt == 7) throw ApiError{SIS_ERR_CAPACITY, "batched lists did not converge in size"};
    // an in-kernel rebuild outgrew a capacity: rebuild the lagging
    // replicas on the per-phase path (which grows every list), then let
    // them finish the segment
    to_packed(c);
    mcf = host_rebuild(c, lagm);
    to_batched(c);
  }
}

// noise amplitude c_r of bead i at temperature T (set_md_coef,
// src/job_md.F90:700-767); the formula upload_conditions applies
double md_cr_at(sis_gpu_ctx *c, double T, int i) {
  if (c->P.bd) return std::sqrt(2.0 * KB_KCAL * T * c->prm.dt / c->fric[i]);
  const double c1 = 0.5 * c->prm.dt * c->fric[i] / c->mass[i];
  const double c2 = std::sqrt(1.0 / (1.0 + c1));
  return 0.5 * c2 / c->mass[i] * std::sqrt(2.0 * c->fric[i] * KB_KCAL * T * c->prm.dt);
}

}  // namespace

// Batched form: one workgroup per replica runs whole segments of steps
// (up to the next exchange).  Lists move to the batched layout for the
// segment and back afterwards, so skin counters, pending rebuilds and the
// limit-MC state carry over exactly as in the per-phase form.
void batched_steps(sis_gpu_ctx *c, int64_t nsteps) {
  if (!batched_ok(c)) {
    for (int64_t s = 0; s < nsteps; s++) step_phases(c);
    return;
  }
  const int R = c->R;
  std::vector<int> all(R, 1), mcf(R, 0);
  if (!c->lists_valid) mcf = host_rebuild(c, all);
  ensure_batched_io(c);
  c->block = batched_block(c);
  int64_t left = nsteps;
  const int ex = c->opt.exchange_every;
  while (left > 0) {
    int64_t seg = left;
    if (ex > 0) seg = std::min<int64_t>(left, ex - (c->istep % ex));
    to_batched(c);
    batched_segment(c, seg, std::vector<int>(R, (int)seg), std::vector<int64_t>(R, c->istep), mcf);
    to_packed(c);
    std::fill(mcf.begin(), mcf.end(), 0);
    c->istep += seg;
    left -= seg;
  }
}

// Whether whole batched segments can run with the exchange on the device:
// a temperature or ionic-strength ladder of the batched form, nothing
// else happening at segment ends.
bool device_exchange_ok(sis_gpu_ctx *c) {
  return batched_ok(c) && c->opt.exchange_every > 0 && c->R > 1 && !(c->opt.flags & SIS_OPT_HOST_EXCHANGE) &&
         !c->anneal_on && !c->ext_on && !c->rg_remd && !c->dcf_remd;
}

// nseg whole segments of the batched form with energy_replica and
// replica_exchange on the device (sis_exchange.h): one segment = the
// steps launch, the counterpart and own-condition energy launches, the
// Metropolis sweep and the rescale, captured once as a HIP graph and
// replayed nseg times with no host synchronisation in between.  Returns
// the steps done (nseg * cadence).  If an in-kernel rebuild outgrows a
// capacity, the rest of the graph does nothing; the host finishes that
// segment and its exchange on its own path and returns early.
int64_t batched_device_run(sis_gpu_ctx *c, int64_t nseg) {
  const int N = c->N, R = c->R, ex = c->opt.exchange_every;
  const int B = batched_block(c);
  const int lpb = lanes_per_bead(N);
  std::vector<int> all(R, 1), mcf(R, 0);
  if (!c->lists_valid) mcf = host_rebuild(c, all);
  ensure_batched_io(c);
  c->block = B;
  // label tables from the host's own formulas
  auto need = [](auto &b, size_t n) { if (b.n < n) b.alloc(n); };
  need(c->dx_rep2lab_d, R + 1); need(c->dx_lab2rep_d, R + 1); need(c->dx_pair_tb_d, c->pair_tb.size());
  need(c->dx_idx_d, 1); need(c->dx_has_d, R); need(c->dx_vsdiv_d, R); need(c->dx_nexch_d, 1);
  need(c->dx_acc_d, R / 2 + 1);
  need(c->dx_labT_d, R + 1); need(c->dx_labCoef_d, R + 1); need(c->dx_labLam_d, R + 1);
  need(c->dx_labCr_d, (size_t)(R + 1) * N); need(c->dx_oppT_d, R); need(c->dx_oppCoef_d, R);
  need(c->dx_oppLam_d, R); need(c->dx_energy_d, (size_t)R * 32); need(c->dx_vsf_d, R);
  std::vector<double> lT(R + 1, 0.0), lC(R + 1, 0.0), lL(R + 1, 1.0), lCr((size_t)(R + 1) * N, 0.0);
  for (int L = 1; L <= R; L++) {
    double cut;
    lT[L] = c->lab2val[L];
    host_set_ele(c->prm, lT[L], c->lab2ion[L], lC[L], lL[L], cut);
    for (int i = 0; i < N; i++) lCr[(size_t)L * N + i] = md_cr_at(c, lT[L], i);
  }
  const int idx0 = c->idx_pair, zero = 0;
  HIPCHK(hipMemcpyAsync(c->dx_labT_d.p, lT.data(), lT.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_labCoef_d.p, lC.data(), lC.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_labLam_d.p, lL.data(), lL.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_labCr_d.p, lCr.data(), lCr.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_rep2lab_d.p, c->rep2lab.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_lab2rep_d.p, c->lab2rep.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_pair_tb_d.p, c->pair_tb.data(), c->pair_tb.size() * sizeof(int), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_idx_d.p, &idx0, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dx_nexch_d.p, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
  to_batched(c);
  std::vector<int> todo(R, ex);
  std::vector<int64_t> stepno(R, c->istep);
  HIPCHK(hipMemcpyAsync(c->bb_todo_d.p, todo.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->bb_stepno_d.p, stepno.data(), R * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->bb_mcfirst_d.p, mcf.data(), R * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->bb_nreb_d.p, 0, R * sizeof(int), c->stream));
  HIPCHK(hipMemsetAsync(c->flags_d.p, 0, 8 * sizeof(int), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));   // the host tables above are released after this
  ExchDev X{};
  X.R = R; X.N = N;
  X.rep2lab = c->dx_rep2lab_d.p; X.lab2rep = c->dx_lab2rep_d.p; X.pair_tb = c->dx_pair_tb_d.p;
  X.idx_pair = c->dx_idx_d.p; X.pair_sets = c->pair_sets;
  X.lab_T = c->dx_labT_d.p; X.lab_coef = c->dx_labCoef_d.p; X.lab_lam = c->dx_labLam_d.p; X.lab_cr = c->dx_labCr_d.p;
  X.opp_T = c->dx_oppT_d.p; X.opp_coef = c->dx_oppCoef_d.p; X.opp_lam = c->dx_oppLam_d.p; X.has_opp = c->dx_has_d.p;
  X.energy = c->dx_energy_d.p; X.vs_f = c->dx_vsf_d.p; X.vs_div = c->dx_vsdiv_d.p;
  X.stepno = c->bb_stepno_d.p; X.accepted = c->dx_acc_d.p; X.nexch = c->dx_nexch_d.p;
  X.seed = c->seed; X.flags = c->flags_d.p;
  // one segment, captured (no memset or host copy inside: the per-segment
  // state lives on the device)
  const void *kfn = batched_kernel(c, B);
  Params P = c->P;
  BatchArgs As = batched_args(c, lpb);
  As.write_back = 1;
  BatchArgs Ao = As, Aw = As;
  Ao.write_back = 0; Ao.energy_mode = 1; Ao.energy_slot = 1; Ao.energy = c->dx_energy_d.p;
  Ao.e_coef = c->dx_oppCoef_d.p; Ao.e_lam = c->dx_oppLam_d.p; Ao.e_use = c->dx_has_d.p;
  Aw.write_back = 0; Aw.energy_mode = 1; Aw.energy_slot = 0; Aw.energy = c->dx_energy_d.p;
  Ptrs Db = batched_ptrs(c), Dopp = Db, Dp = c->ptrs();
  Dopp.tK = c->dx_oppT_d.p;   // the energies at the partner's temperature
  const size_t lds = batched_lds_bytes(As.xlds, As.fplds, As.nblds, N, As.stlds, As.mplds);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  {
    void *a1[] = {&P, &Db, &As};
    HIPCHK(hipLaunchKernel(kfn, dim3(R), dim3(B), a1, lds, c->stream));
    hipLaunchKernelGGL(k_exchange_prep, dim3((R + 255) / 256), dim3(256), 0, c->stream, X, Dp);
    void *a2[] = {&P, &Dopp, &Ao};
    HIPCHK(hipLaunchKernel(kfn, dim3(R), dim3(B), a2, lds, c->stream));
    void *a3[] = {&P, &Db, &Aw};
    HIPCHK(hipLaunchKernel(kfn, dim3(R), dim3(B), a3, lds, c->stream));
    hipLaunchKernelGGL(k_exchange_decide, dim3(1), dim3(64), 0, c->stream, X);
    hipLaunchKernelGGL(k_exchange_apply, dim3((N + 255) / 256, R), dim3(256), 0, c->stream, X, Dp);
  }
  HIPCHK(hipStreamEndCapture(c->stream, &graph));
  HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  HIPCHK(hipGraphDestroy(graph));
  for (int64_t sgi = 0; sgi < nseg; sgi++) {
    c->tmark();
    HIPCHK(hipGraphLaunch(exec, c->stream));
    c->tmark();
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipGraphExecDestroy(exec));
  std::vector<int> f(8), nr(R), lab(R + 1), l2r(R + 1);
  int idx = 0, nexch = 0;
  HIPCHK(hipMemcpy(f.data(), c->flags_d.p, 8 * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nr.data(), c->bb_nreb_d.p, R * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(lab.data(), c->dx_rep2lab_d.p, (R + 1) * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(l2r.data(), c->dx_lab2rep_d.p, (R + 1) * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&idx, c->dx_idx_d.p, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&nexch, c->dx_nexch_d.p, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(stepno.data(), c->bb_stepno_d.p, R * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (f[7]) throw ApiError{SIS_ERR_STATE, "limit MC met a base-pair candidate outside the system"};
  if (f[1] == 2) throw ApiError{SIS_ERR_STATE, "limit MC found an excess nucleotide without pairs"};
  if (f[5]) throw ApiError{SIS_ERR_CAPACITY, "a base-pair conflict component outgrew the limit-MC scratch"};
  for (int r = 0; r < R; r++) c->rebuilds += nr[r];
  // the host's copies of the labels and the conditions they carry
  for (int k = 0; k <= R; k++) { c->rep2lab[k] = lab[k]; c->lab2rep[k] = l2r[k]; }
  for (int r = 0; r < R; r++) {
    c->tK[r] = c->lab2val[c->rep2lab[r + 1]];
    c->ionic[r] = c->lab2ion[c->rep2lab[r + 1]];
  }
  c->idx_pair = idx;
  c->dev_exchanges += nexch;
  const int64_t istep0 = c->istep;
  if (!f[0]) {
    to_packed(c);
    c->istep = istep0 + nseg * ex;
    return nseg * ex;
  }
  // a capacity overflow in segment `nexch`: every later launch did
  // nothing.  Finish that segment on the host path, then its exchange.
  const int64_t seg_end = istep0 + (int64_t)(nexch + 1) * ex;
  std::vector<int> lagm(R, 0);
  for (int r = 0; r < R; r++) {
    todo[r] = (int)(seg_end - stepno[r]);
    lagm[r] = todo[r] > 0 ? 1 : 0;
  }
  HIPCHK(hipMemset(c->flags_d.p, 0, 8 * sizeof(int)));
  to_packed(c);
  mcf = host_rebuild(c, lagm);
  to_batched(c);
  batched_segment(c, ex, todo, stepno, mcf);
  to_packed(c);
  c->istep = seg_end;
  do_exchange(c, nullptr, nullptr);
  return (int64_t)(nexch + 1) * ex;
}

namespace {

// neighbour-list skin of the context (neighbor_list_margin): the list
// cutoffs and the cell grid follow it
void apply_margin(sis_gpu_ctx *c, double m) {
  c->prm.nl_margin = m;
  c->P.nl_margin = m;
  c->P.wca_nl_cut2 = (c->P.wca_sigma + m) * (c->P.wca_sigma + m);
  if (c->P.bp_model != 0) c->P.bp_nl_cut2 = (c->P.bp_cutoff_dist + m) * (c->P.bp_cutoff_dist + m);
  setup_cells(c);
  c->lists_valid = false;
}

}  // namespace

// Autotuner of the launch shape and the rebuild skin: from the current
// state, times nsteps steps of the context's form at every candidate
// (batched: threads per replica 128 / 256 / 512; persistent: 1, 2 and as
// many workgroups per compute unit as fit; per-phase steps in batches:
// 0 replayed from a captured graph, 1 issued directly) times every skin,
// restoring the state before each trial and at the end, and keeps the
// fastest (batched: the narrowest shape within 3% of it).
// The skin changes when lists are rebuilt, not the forces (every term
// applies its own cutoff), so the choice is a matter of speed only.
void autotune(sis_gpu_ctx *c, const double *skins, int nskin, int64_t nsteps, double *report) {
  int mode = c->mode ? c->mode : choose_mode(c);
  const bool batched = mode == 3 && batched_ok(c);
  const bool pers = mode == 2 && c->R == 1 && !c->ext_on;
  const bool phases = !batched && !pers && mode == 1 && graph_steps_ok(c);
  if (!batched && !pers && !phases)
    throw ApiError{SIS_ERR_UNSUPPORTED, "the autotuner covers the batched, persistent and per-phase graph forms"};
  const int R = c->R, N = c->N;
  const size_t RN = (size_t)R * N;
  // saved state: coordinates, velocities, noise terms, step, labels
  std::vector<double> x(3 * RN), v(3 * RN), a(3 * RN);
  {
    std::vector<double4> h(RN);
    auto down = [&](std::vector<double> &dst, DevBuf<double4> &src) {
      HIPCHK(hipMemcpy(h.data(), src.p, RN * sizeof(double4), hipMemcpyDeviceToHost));
      for (size_t g = 0; g < RN; g++) { dst[3 * g] = h[g].x; dst[3 * g + 1] = h[g].y; dst[3 * g + 2] = h[g].z; }
    };
    down(x, c->pos_d); down(v, c->vel_d); down(a, c->acc_d);
  }
  const int64_t istep0 = c->istep, reb0 = c->rebuilds, dex0 = c->dev_exchanges;
  const std::vector<int> lab0(c->rep2lab.begin() + 1, c->rep2lab.end());
  const int idx0 = c->idx_pair;
  const double m0 = c->prm.nl_margin;
  auto restore = [&]() {
    if (R > 1) { std::vector<int32_t> l(lab0.begin(), lab0.end()); sis_gpu_set_labels(c, l.data()); }
    c->idx_pair = idx0;
    sis_gpu_set_state(c, x.data(), v.data(), a.data());
    c->istep = istep0;
    c->lists_valid = false;
  };
  std::vector<int> shapes;
  if (batched) {
    shapes = {128, 256, 512};
  } else if (phases) {
    shapes = {0, 1};   // batches of steps replayed from a graph, or issued directly
  } else {
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    shapes = {ncu, 2 * ncu, 0};
  }
  double best = 1e300;
  int best_shape = batched ? 256 : 0;
  double best_skin = m0;
  // per batched shape: its fastest skin
  std::vector<double> shape_ms(shapes.size(), 1e300);
  std::vector<double> shape_skin(shapes.size(), m0);
  int k = 0;
  for (size_t si = 0; si < shapes.size(); si++) {
    const int sh = shapes[si];
    for (int q = 0; q < nskin; q++) {
      if (!(skins[q] > 0.0)) throw ApiError{SIS_ERR_ARG, "skin must be positive"};
      if (batched) c->opt.block = sh; else if (pers) c->opt.grid = sh; else c->pg_eager = sh == 1;
      apply_margin(c, skins[q]);
      restore();
      HIPCHK(hipStreamSynchronize(c->stream));
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      int64_t rb = 0;
      int rc = sis_gpu_step(c, nsteps, &rb);
      if (rc != SIS_OK) throw ApiError{rc, std::string("autotune trial failed: ") + g_err};
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      HIPCHK(hipEventSynchronize(c->ev1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      const double per = ms / (double)nsteps;
      if (report) {
        report[2 + 3 * k] = batched ? sh : (pers ? c->block : sh);
        report[3 + 3 * k] = skins[q];
        report[4 + 3 * k] = per;
      }
      k++;
      if (per < best) { best = per; best_shape = sh; best_skin = skins[q]; }
      if (per < shape_ms[si]) { shape_ms[si] = per; shape_skin[si] = skins[q]; }
    }
  }
  // Batched form: the narrowest workgroup within 3% of the fastest trial.
  // Trials of a few hundred steps differ by about that much from run to run,
  // and a 512-thread workgroup runs two waves per SIMD, whose 256-register
  // cap makes k_batched spill to scratch (HBM writes) where 256 threads do not.
  if (batched)
    for (size_t si = 0; si < shapes.size(); si++)
      if (shape_ms[si] <= best * 1.03) {
        best_shape = shapes[si];
        best_skin = shape_skin[si];
        break;
      }
  if (batched) c->opt.block = best_shape; else if (pers) c->opt.grid = best_shape; else c->pg_eager = best_shape == 1;
  apply_margin(c, best_skin);
  restore();
  c->rebuilds = reb0;
  c->dev_exchanges = dex0;
  if (report) { report[0] = best_shape; report[1] = best_skin; }
}
