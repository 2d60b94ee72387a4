// Job pool: the serving form.  Independent MD jobs (own topology,
// temperature, salt, box, step count and seed; shared force field) are
// submitted as they arrive.  A scheduler thread admits them in arrival
// order while their HBM footprint fits the budget, keeps each job's state,
// lists and random-stream position resident in its own single-replica
// context, and advances every resident job each round:
//   - jobs of up to POOL_WG_MAX beads in shared launches of k_jobs (one
//     workgroup per job, one launch per precision class);
//   - larger jobs, which one compute unit would serve too slowly, on the
//     per-phase path over the whole GPU, several at once from worker
//     threads, overlapping the shared launches.
// Each round gives a job about `quantum` bead-steps, so a round takes
// about the same time for every job size and new arrivals join at the
// next round.  A job's trajectory depends only on its own inputs: its path
// follows from its size, its RNG streams are keyed by its seed, its step
// chunks by its size, and rounds continue it exactly (lists, pending
// rebuilds and MC state carry over).
// Included by sis_gpu.hip after the batched form's driver.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

constexpr int POOL_WG_MAX = 2048;     // largest job run by one workgroup
constexpr int POOL_LARGE_PAR = 8;     // large jobs stepped at once

struct PoolJob {
  int64_t id = 0;
  std::vector<int32_t> nmp_chain, moltype, seq;
  sis_params prm{};
  std::vector<double> xyz;
  int precision = 0;
  int64_t nsteps = 0;
  uint64_t seed = 0;
  int nmp = 0;
  // scheduler state
  sis_gpu_ctx *c = nullptr;
  int slot = -1;
  int64_t done = 0, rebuilds = 0, todo = 0;
  int mc_first = 1;
  size_t bytes = 0;
  // reported (under the pool mutex)
  int state = SIS_JOB_QUEUED;
  double t_submit = 0, t_start = -1, t_end = -1;
  std::string err;
  std::vector<double> result;
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct sis_pool {
  int device = 0;
  uint64_t budget = 0;
  int64_t quantum = 0;
  int max_active = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<PoolJob>> arrivals;   // submitted, not yet taken by the scheduler
  std::unordered_map<int64_t, std::shared_ptr<PoolJob>> jobs;
  int64_t next_id = 1;
  bool stop = false;
  std::thread th;
  // scheduler-owned
  std::deque<std::shared_ptr<PoolJob>> pending;
  std::vector<std::shared_ptr<PoolJob>> active;
  std::vector<int> free_slots;
  size_t bytes_used = 0;
  bool have_P0 = false;
  Params P0{};
  BpPara bpp0[7] = {};
  hipStream_t stream[2] = {nullptr, nullptr};
  DevBuf<int> io_d, flags_d;
  DevBuf<JobView> views_d[2];
  // statistics (under mu)
  double bead_steps = 0, gpu_s = 0;
  int64_t launches = 0, completed = 0, failed = 0;
};

namespace {

// sizes compared between jobs of one pool: everything but the system size
bool same_force_field(Params a, Params b) {
  a.N = b.N = 0;
  a.R = b.R = 0;
  a.nchains = b.nchains = 0;
  a.maxlen = b.maxlen = 0;
  return std::memcmp(&a, &b, sizeof(Params)) == 0;
}

size_t job_estimate(const PoolJob &j) { return (size_t)j.nmp * 4096 + ((size_t)8 << 20); }

void job_fail(sis_pool *p, PoolJob &j, const std::string &msg) {
  std::lock_guard<std::mutex> lk(p->mu);
  j.state = SIS_JOB_FAILED;
  j.err = msg;
  j.t_end = now_s(p->t0);
  p->failed++;
}

void job_release(sis_pool *p, PoolJob &j) {
  if (j.c) {
    sis_gpu_destroy(j.c);
    j.c = nullptr;
  }
  if (j.slot >= 0) p->free_slots.push_back(j.slot);
  j.slot = -1;
  p->bytes_used -= std::min(p->bytes_used, j.bytes);
  j.bytes = 0;
}

// context of a job: set up on the per-phase path (lists, velocities)
bool job_admit(sis_pool *p, PoolJob &j) {
  size_t f0 = 0, f1 = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&f0, &tot));
  sis_topo t{};
  t.nchains = (int32_t)j.nmp_chain.size();
  t.nmp = j.nmp;
  t.nmp_chain = j.nmp_chain.data();
  t.moltype = j.moltype.data();
  t.seq = j.seq.data();
  sis_options o{};
  o.nrep = 1;
  o.precision = j.precision;
  o.mode = 1;
  o.seed = j.seed;
  // the pool's worker threads launch steps while its scheduler thread
  // allocates and frees other jobs' memory: no stream capture there
  o.flags = SIS_OPT_NO_GRAPH;
  sis_gpu_ctx *c = nullptr;
  if (sis_gpu_create(&t, &j.prm, &o, &c) != SIS_OK) { job_fail(p, j, g_err); return false; }
  j.c = c;
  if (!p->have_P0) {
    p->P0 = c->P;
    std::memcpy(p->bpp0, c->bpp_h, sizeof p->bpp0);
    p->have_P0 = true;
  }
  if (!same_force_field(p->P0, c->P) || std::memcmp(p->bpp0, c->bpp_h, sizeof p->bpp0) != 0) {
    job_fail(p, j, "job force field differs from the pool's");
    job_release(p, j);
    return false;
  }
  if (sis_gpu_set_state(c, j.xyz.data(), nullptr, nullptr) != SIS_OK || sis_gpu_rebuild(c, 0) != SIS_OK ||
      sis_gpu_init_velocities(c) != SIS_OK) {
    job_fail(p, j, g_err);
    job_release(p, j);
    return false;
  }
  if (c->seg_d.n < (size_t)c->N * SEGCAP) c->seg_d.alloc((size_t)c->N * SEGCAP);
  j.slot = p->free_slots.back();
  p->free_slots.pop_back();
  int io[4] = {0, c->nbp_total, 0, 0};
  HIPCHK(hipMemcpy(p->io_d.p + 4 * j.slot, io, sizeof io, hipMemcpyHostToDevice));
  HIPCHK(hipMemGetInfo(&f1, &tot));
  j.bytes = f0 > f1 ? f0 - f1 : job_estimate(j);
  p->bytes_used += j.bytes;
  j.mc_first = c->P.nstep_bp_mc > 0 ? 1 : 0;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    j.state = SIS_JOB_RUNNING;
    j.t_start = now_s(p->t0);
  }
  return true;
}

JobView job_view(sis_pool *p, PoolJob &j) {
  sis_gpu_ctx *c = j.c;
  const Ptrs d = c->ptrs();
  JobView v{};
  v.N = c->N;
  v.nchains = c->P.nchains;
  v.maxlen = c->P.maxlen;
  v.capb = c->D.maxbp;
  v.lpb = lanes_per_bead(c->N);
  v.rk = 0;
  v.todo = (int)j.todo;
  v.mc_first = j.mc_first;
  v.maxnb = d.maxnb;
  v.maxw = d.maxw;
  v.maxinc = d.maxinc;
  v.maxbpb = d.maxbpb;
  v.stepno = c->istep;
  v.seed = c->seed;
  v.topo = d.topo; v.chain_id = d.chain_id; v.bp3_ctx = d.bp3_ctx;
  v.md_a = d.md_a; v.md_cf = d.md_cf; v.md_cx = d.md_cx; v.md_cr = d.md_cr;
  v.tK = d.tK; v.ele_coef = d.ele_coef; v.lambdaD = d.lambdaD; v.ele_cut = d.ele_cut; v.box = d.box;
  v.pos = d.pos; v.pos_b = c->pos2_d.p; v.vel = d.vel; v.acc = d.acc; v.move = d.move; v.frc = d.frc;
  v.fbond = d.fbond;
  v.nb_ele = d.nb_ele; v.n_ele = d.n_ele; v.nb_wca = d.nb_wca; v.n_wca = d.n_wca;
  v.bp_tmp = d.bp_tmp; v.n_bp_tmp = d.n_bp_tmp; v.bp_off = d.bp_off;
  v.bp_mp = d.bp_mp; v.bp_coef = d.bp_coef; v.inc = d.inc; v.n_inc = d.n_inc;
  v.bp_status = d.bp_status; v.bp_status_mc = d.bp_status_mc; v.ene_bp = d.ene_bp; v.for_bp = d.for_bp;
  v.nt_cnt = d.nt_cnt; v.comp = d.comp;
  v.flags = p->flags_d.p + 8 * j.slot;
  v.seg_n = c->exc_cnt_d.p; v.seg = c->seg_d.p; v.has_exc = c->exc_fill_d.p;
  v.io = p->io_d.p + 4 * j.slot;
  return v;
}

// one round: every resident job advances by its chunk
void pool_round(sis_pool *p) {
  std::vector<JobView> views[2];
  std::vector<PoolJob *> members[2], large;
  for (auto &jp : p->active) {
    PoolJob &j = *jp;
    j.todo = std::min<int64_t>(j.nsteps - j.done, std::max<int64_t>(1, p->quantum / std::max(j.nmp, 1)));
    if (j.nmp > POOL_WG_MAX) {
      large.push_back(&j);
      continue;
    }
    const int k = j.precision == 1 ? 1 : 0;
    views[k].push_back(job_view(p, j));
    members[k].push_back(&j);
  }
  HIPCHK(hipMemsetAsync(p->flags_d.p, 0, p->flags_d.n * sizeof(int), p->stream[0]));
  HIPCHK(hipStreamSynchronize(p->stream[0]));
  auto t0 = std::chrono::steady_clock::now();
  int nl = 0;
  for (int k = 0; k < 2; k++) {
    if (views[k].empty()) continue;
    HIPCHK(hipMemcpyAsync(p->views_d[k].p, views[k].data(), views[k].size() * sizeof(JobView),
                          hipMemcpyHostToDevice, p->stream[k]));
    const Ptrs D0 = members[k].front()->c->ptrs();   // the shared base-pair tables
    // dynamic LDS for the positions of the largest member that fits
    int xlds = 0, fplds = 0;
    for (PoolJob *jp : members[k])
      if (jp->nmp <= BATCH_X_LDS) {
        xlds = std::max(xlds, jp->nmp);
        const int fp = jp->nmp * lanes_per_bead(jp->nmp);
        if (fp <= BATCH_FP_LDS) fplds = std::max(fplds, fp);
      }
    const size_t dyn = (size_t)(xlds + fplds) * sizeof(double4);
    if (k == 1)
      hipLaunchKernelGGL((k_jobs<float, 256>), dim3((unsigned)views[k].size()), dim3(256), dyn, p->stream[k], p->P0,
                         D0, (const JobView *)p->views_d[k].p, xlds, fplds);
    else
      hipLaunchKernelGGL((k_jobs<double, 256>), dim3((unsigned)views[k].size()), dim3(256), dyn, p->stream[k],
                         p->P0, D0, (const JobView *)p->views_d[k].p, xlds, fplds);
    HIPCHK(hipGetLastError());
    nl++;
  }
  // large jobs: whole-GPU per-phase steps, POOL_LARGE_PAR at a time, while
  // the shared launches run
  std::vector<std::string> lerr(large.size());
  std::vector<int64_t> lreb(large.size(), 0);
  for (size_t b = 0; b < large.size(); b += POOL_LARGE_PAR) {
    std::vector<std::thread> th;
    for (size_t i = b; i < std::min(large.size(), b + POOL_LARGE_PAR); i++)
      th.emplace_back([p, &large, &lerr, &lreb, i]() {
        if (hipSetDevice(p->device) != hipSuccess) { lerr[i] = "hipSetDevice failed"; return; }
        if (sis_gpu_step(large[i]->c, large[i]->todo, &lreb[i]) != SIS_OK) lerr[i] = sis_gpu_last_error();
      });
    for (auto &t : th) t.join();
    nl++;
  }
  for (int k = 0; k < 2; k++) HIPCHK(hipStreamSynchronize(p->stream[k]));
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::vector<int> io(p->io_d.n), fl(p->flags_d.n);
  HIPCHK(hipMemcpy(io.data(), p->io_d.p, io.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(fl.data(), p->flags_d.p, fl.size() * sizeof(int), hipMemcpyDeviceToHost));
  double bs = 0;
  for (int k = 0; k < 2; k++)
    for (PoolJob *jp : members[k]) {
      PoolJob &j = *jp;
      const int *jio = io.data() + 4 * j.slot, *f = fl.data() + 8 * j.slot;
      const int d = jio[0];
      j.done += d;
      j.c->istep += d;
      j.rebuilds += jio[3];
      bs += (double)d * j.nmp;
      j.mc_first = 0;
      if (f[7]) { job_fail(p, j, "limit MC met a base-pair candidate outside the system"); continue; }
      if (f[1] == 2) { job_fail(p, j, "limit MC found an excess nucleotide without pairs"); continue; }
      if (f[5]) { job_fail(p, j, "a base-pair conflict component outgrew the limit-MC scratch"); continue; }
      if (d < j.todo) {
        // an in-kernel rebuild outgrew a capacity: rebuild on the host,
        // which grows the lists, and carry on next round
        try {
          std::vector<int> all(1, 1);
          rebuild(j.c, all);
          if ((size_t)j.c->N * SEGCAP > j.c->seg_d.n) j.c->seg_d.alloc((size_t)j.c->N * SEGCAP);
          j.rebuilds++;
          j.mc_first = j.c->P.nstep_bp_mc > 0 ? 1 : 0;
          int nio[4] = {0, j.c->nbp_total, 0, 0};
          HIPCHK(hipMemcpy(p->io_d.p + 4 * j.slot, nio, sizeof nio, hipMemcpyHostToDevice));
        } catch (const ApiError &e) {
          job_fail(p, j, e.msg);
        }
      }
    }
  for (size_t i = 0; i < large.size(); i++) {
    PoolJob &j = *large[i];
    if (!lerr[i].empty()) { job_fail(p, j, lerr[i]); continue; }
    j.done += j.todo;
    j.rebuilds += lreb[i];
    bs += (double)j.todo * j.nmp;
  }
  std::lock_guard<std::mutex> lk(p->mu);
  p->bead_steps += bs;
  p->gpu_s += dt;
  p->launches += nl;
}

// finished and failed jobs leave the device
void pool_retire(sis_pool *p) {
  std::vector<std::shared_ptr<PoolJob>> keep;
  for (auto &jp : p->active) {
    PoolJob &j = *jp;
    int st;
    {
      std::lock_guard<std::mutex> lk(p->mu);
      st = j.state;
    }
    if (st == SIS_JOB_RUNNING && j.done < j.nsteps) {
      keep.push_back(jp);
      continue;
    }
    if (st == SIS_JOB_RUNNING) {
      std::vector<double> x((size_t)3 * j.nmp);
      const bool ok = sis_gpu_get_state(j.c, x.data(), nullptr, nullptr) == SIS_OK;
      std::lock_guard<std::mutex> lk(p->mu);
      if (ok) {
        j.result.swap(x);
        j.state = SIS_JOB_DONE;
        p->completed++;
      } else {
        j.state = SIS_JOB_FAILED;
        j.err = g_err;
        p->failed++;
      }
      j.t_end = now_s(p->t0);
    }
    job_release(p, j);
  }
  p->active.swap(keep);
}

void pool_admit(sis_pool *p) {
  while (!p->pending.empty() && (int)p->active.size() < p->max_active && !p->free_slots.empty()) {
    auto jp = p->pending.front();
    if (!p->active.empty() && p->bytes_used + job_estimate(*jp) > p->budget) break;   // wait for memory
    p->pending.pop_front();
    if (job_admit(p, *jp)) p->active.push_back(jp);
  }
}

void pool_loop(sis_pool *p) {
  if (hipSetDevice(p->device) != hipSuccess) return;
  while (true) {
    {
      std::unique_lock<std::mutex> lk(p->mu);
      if (p->stop) break;
      while (!p->arrivals.empty()) {
        p->pending.push_back(p->arrivals.front());
        p->arrivals.pop_front();
      }
      if (p->pending.empty() && p->active.empty()) {
        p->cv.wait_for(lk, std::chrono::milliseconds(5));
        continue;
      }
    }
    try {
      pool_admit(p);
      if (!p->active.empty()) pool_round(p);
      pool_retire(p);
    } catch (const HipError &e) {
      // a device error ends every resident job
      for (auto &jp : p->active) {
        job_fail(p, *jp, e.msg);
        jp->c = nullptr;   // the context may be unusable; leak it rather than touch the device
      }
      p->active.clear();
    } catch (const ApiError &e) {
      for (auto &jp : p->active) job_fail(p, *jp, e.msg);
      for (auto &jp : p->active) job_release(p, *jp);
      p->active.clear();
    }
  }
  for (auto &jp : p->active) job_release(p, *jp);
  p->active.clear();
}

}  // namespace

extern "C" {

int sis_pool_create(uint64_t hbm_budget_bytes, int64_t quantum_bead_steps, int32_t max_active, sis_pool **out) {
  API_BEGIN
  if (!out) return fail(SIS_ERR_ARG, "null argument");
  *out = nullptr;
  if (max_active <= 0) max_active = 1024;
  if (quantum_bead_steps <= 0) quantum_bead_steps = 2000000;
  auto p = std::make_unique<sis_pool>();
  HIPCHK(hipGetDevice(&p->device));
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  p->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(fr * 0.8);
  p->quantum = quantum_bead_steps;
  p->max_active = max_active;
  p->t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < 2; k++) HIPCHK(hipStreamCreateWithFlags(&p->stream[k], hipStreamNonBlocking));
  p->io_d.alloc((size_t)4 * max_active);
  p->flags_d.alloc((size_t)8 * max_active);
  for (int k = 0; k < 2; k++) p->views_d[k].alloc((size_t)max_active);
  for (int s = max_active - 1; s >= 0; s--) p->free_slots.push_back(s);
  sis_pool *raw = p.release();
  raw->th = std::thread(pool_loop, raw);
  *out = raw;
  return SIS_OK;
  API_END
}

int sis_pool_submit(sis_pool *p, const sis_topo *topo, const sis_params *prm, const double *xyz, int32_t precision,
                    int64_t nsteps, uint64_t seed, int64_t *job_id) {
  API_BEGIN
  if (!p || !topo || !prm || !xyz || !job_id) return fail(SIS_ERR_ARG, "null argument");
  if (topo->nchains <= 0 || topo->nmp <= 0) return fail(SIS_ERR_ARG, "empty topology");
  if (nsteps < 0) return fail(SIS_ERR_ARG, "nsteps < 0");
  if (precision != 0 && precision != 1) return fail(SIS_ERR_ARG, "precision must be 0 or 1");
  auto j = std::make_shared<PoolJob>();
  j->nmp_chain.assign(topo->nmp_chain, topo->nmp_chain + topo->nchains);
  j->moltype.assign(topo->moltype, topo->moltype + topo->nchains);
  j->seq.assign(topo->seq, topo->seq + topo->nmp);
  j->prm = *prm;
  j->xyz.assign(xyz, xyz + (size_t)3 * topo->nmp);
  j->precision = precision;
  j->nsteps = nsteps;
  j->seed = seed;
  j->nmp = topo->nmp;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    j->id = p->next_id++;
    j->t_submit = now_s(p->t0);
    p->jobs[j->id] = j;
    p->arrivals.push_back(j);
  }
  p->cv.notify_one();
  *job_id = j->id;
  return SIS_OK;
  API_END
}

int sis_pool_poll(sis_pool *p, int64_t job_id, int32_t *state, double *times, double *xyz) {
  API_BEGIN
  if (!p || !state) return fail(SIS_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  auto it = p->jobs.find(job_id);
  if (it == p->jobs.end()) return fail(SIS_ERR_ARG, "unknown job id");
  const PoolJob &j = *it->second;
  *state = j.state;
  if (times) {
    times[0] = j.t_submit;
    times[1] = j.t_start;
    times[2] = j.t_end;
    times[3] = (double)j.done;
    times[4] = (double)j.rebuilds;
  }
  if (xyz && j.state == SIS_JOB_DONE) std::copy(j.result.begin(), j.result.end(), xyz);
  if (j.state == SIS_JOB_FAILED) return fail(SIS_ERR_STATE, j.err);
  return SIS_OK;
  API_END
}

int sis_pool_forget(sis_pool *p, int64_t job_id) {
  API_BEGIN
  if (!p) return fail(SIS_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  auto it = p->jobs.find(job_id);
  if (it == p->jobs.end()) return fail(SIS_ERR_ARG, "unknown job id");
  if (it->second->state == SIS_JOB_QUEUED || it->second->state == SIS_JOB_RUNNING)
    return fail(SIS_ERR_STATE, "job still queued or running");
  p->jobs.erase(it);
  return SIS_OK;
  API_END
}

int sis_pool_stats(sis_pool *p, double *out, int32_t n) {
  API_BEGIN
  if (!p || !out) return fail(SIS_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  int64_t queued = 0, running = 0;
  for (auto &kv : p->jobs) {
    queued += kv.second->state == SIS_JOB_QUEUED;
    running += kv.second->state == SIS_JOB_RUNNING;
  }
  const double v[9] = {(double)queued, (double)running, (double)p->completed, (double)p->failed, p->bead_steps,
                       p->gpu_s, (double)p->launches, (double)p->bytes_used, now_s(p->t0)};
  for (int i = 0; i < n && i < 9; i++) out[i] = v[i];
  return SIS_OK;
  API_END
}

int sis_pool_destroy(sis_pool *p) {
  if (!p) return SIS_OK;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    p->stop = true;
  }
  p->cv.notify_one();
  if (p->th.joinable()) p->th.join();
  for (int k = 0; k < 2; k++)
    if (p->stream[k]) (void)hipStreamDestroy(p->stream[k]);
  p->io_d.free_();
  p->flags_d.free_();
  for (int k = 0; k < 2; k++) p->views_d[k].free_();
  delete p;
  return SIS_OK;
}

}  // extern "C"
