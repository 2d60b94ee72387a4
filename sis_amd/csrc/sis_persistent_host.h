// Host drivers of the persistent (mode 2) and batched (mode 3) step forms.
// Included by sis_gpu.hip after the per-phase driver.
#pragma once

int choose_mode(sis_gpu_ctx *c) {
  if (c->parity) return 1;
  if (c->R == 1 && c->N >= 1024) return 2;
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
  if (c->opt.block > 0) g = std::min(cap, std::max(1, c->opt.block));
  return std::max(g, 1);
}

}  // namespace

void persistent_steps(sis_gpu_ctx *c, int64_t nsteps) {
  if (c->R != 1) {
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
    Params P = c->P;
    void *args[] = {&P, &D, &A};
    c->tmark();
    HIPCHK(hipLaunchCooperativeKernel(kfn, dim3(G), dim3(B), args, 0, c->stream));
    c->tmark();
    int done[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(done, c->pers_sync_d.p, sizeof done, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
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

void batched_steps(sis_gpu_ctx *c, int64_t nsteps) {
  for (int64_t s = 0; s < nsteps; s++) step_phases(c);
}
