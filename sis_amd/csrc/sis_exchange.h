// On-device replica exchange for the batched form: energy_replica and
// replica_exchange (src/energy_replica.F90:47-156,
// src/replica_exchange.F90:35-154, src/job_md.F90:536-567) over labels
// that carry a temperature and an ionic strength (ladders of either, or a
// grid of both), with no host round trip.  The conditions every label carries (temperature, DH coefficient
// and Debye length, the noise amplitudes md_cr) are tabulated by the host
// once, from the same formulas the host path evaluates per exchange, so
// both paths give the same numbers.  Included by sis_gpu.hip only.
#pragma once

namespace sis {

struct ExchDev {
  int R, N;
  int *rep2lab, *lab2rep;     // [R+1], 1-based as var_replica's tables
  const int *pair_tb;         // [pair_sets][R+2]
  int *idx_pair;              // [1] which set of the pair table is next
  int pair_sets;              // 2 x the number of replica dimensions
  const double *lab_T, *lab_coef, *lab_lam;   // [R+1]
  const double *lab_cr;       // [(R+1) * N]
  double *opp_T, *opp_coef, *opp_lam;          // [R] counterpart conditions of this exchange
  int *has_opp;               // [R]
  double *energy;             // [R][2][16]: own, counterpart energies
  double *vs_f;               // [R] velocity factor of this exchange (1: unchanged)
  int *vs_div;                // [R] 1: divide by vs_f (src/replica_exchange.F90:175-213)
  int64_t *stepno;            // [R] istep of every replica
  int *accepted;              // [R/2 + 1] per pair, or null
  int *nexch;                 // [1] exchanges done on the device
  uint64_t seed;
  const int *flags;           // the context's error flags: [0] set = capacity overflow earlier
};

// counterpart conditions of every replica for this exchange
__global__ void k_exchange_prep(ExchDev X, Ptrs D) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= X.R || X.flags[0]) return;
  const int lab = X.rep2lab[r + 1];
  const int opp = X.pair_tb[X.idx_pair[0] * (X.R + 2) + lab];
  const bool has = opp > 0 && opp <= X.R;
  X.has_opp[r] = has ? 1 : 0;
  X.opp_T[r] = has ? X.lab_T[opp] : D.tK[r];
  X.opp_coef[r] = has ? X.lab_coef[opp] : D.ele_coef[r];
  X.opp_lam[r] = has ? X.lab_lam[opp] : D.lambdaD[r];
}

// Metropolis sweep over the pairs of this half of the table in label
// order (Sugita & Okamoto, eq. 14 of src/replica_exchange.F90:111-114);
// uniform k of the exchange is Philox keyed by the step, as the host path
// draws it.  One thread: a few hundred pairs at most.
__global__ void k_exchange_decide(ExchDev X) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || X.flags[0]) return;
  const int R = X.R;
  const int64_t istep = X.stepno[0];
  const double KB = 1.380649e-23 * ((1.0 / 4184.0) * 6.02214076e23);
  for (int r = 0; r < R; r++) { X.vs_f[r] = 1.0; X.vs_div[r] = 0; }
  const int *tb = X.pair_tb + X.idx_pair[0] * (R + 2);
  int nr = 0;
  for (int li = 1; li <= R; li++) {
    const int lj = tb[li];
    if (lj < li) continue;
    const int k = nr++;
    const int ri = X.lab2rep[li], rj = X.lab2rep[lj];
    const double tm = X.lab_T[li], tn = X.lab_T[lj];
    const double pim = X.energy[((size_t)(ri - 1) * 2) * 16], pin = X.energy[((size_t)(ri - 1) * 2 + 1) * 16];
    const double pjn = X.energy[((size_t)(rj - 1) * 2) * 16], pjm = X.energy[((size_t)(rj - 1) * 2 + 1) * 16];
    const double delta = ((pjm - pim) / tm - (pjn - pin) / tn) / KB;
    const U4 o = philox(U4{(uint32_t)istep, (uint32_t)(istep >> 32), (uint32_t)k, 0x72657863u}, (uint32_t)X.seed,
                        (uint32_t)(X.seed >> 32) ^ 0x1234567u);
    const double u = u01(o.x, o.y);
    const bool acc = (delta < 0.0) || (u <= exp(-delta));
    if (X.accepted) X.accepted[k] = acc ? 1 : 0;
    if (acc) {
      X.rep2lab[ri] = lj; X.rep2lab[rj] = li;
      X.lab2rep[li] = rj; X.lab2rep[lj] = ri;
      const double sf = sqrt(tn / tm);
      X.vs_f[ri - 1] = sf; X.vs_div[ri - 1] = 0;
      X.vs_f[rj - 1] = sf; X.vs_div[rj - 1] = 1;
    }
  }
  X.idx_pair[0] = (X.idx_pair[0] + 1) % X.pair_sets;
  X.nexch[0] += 1;
}

// velocities rescaled and every replica given its label's conditions
__global__ void k_exchange_apply(ExchDev X, Ptrs D) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (a >= X.N || X.flags[0]) return;
  const size_t g = (size_t)r * X.N + a;
  const double f = X.vs_f[r];
  double4 &v = D.vel[g];
  if (X.vs_div[r]) { v.x = v.x / f; v.y = v.y / f; v.z = v.z / f; }
  else if (f != 1.0) { v.x = f * v.x; v.y = f * v.y; v.z = f * v.z; }
  const int lab = X.rep2lab[r + 1];
  D.md_cr[g] = X.lab_cr[(size_t)lab * X.N + a];
  if (a == 0) {
    D.tK[r] = X.lab_T[lab];
    D.ele_coef[r] = X.lab_coef[lab];
    D.lambdaD[r] = X.lab_lam[lab];
  }
}

}  // namespace sis
