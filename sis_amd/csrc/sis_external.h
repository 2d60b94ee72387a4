// External terms beside the bonded and pair potentials: the wall at z = 0
// (force_stage / energy_stage), optical tweezers (constant-force pairs and
// force ramps; force_tweezers / energy_tweezers), the radius-of-gyration
// bias (force_rg / energy_rg) and sigmoid restraints (force_restraint /
// energy_restraint); src/force.F90:62-75, src/energy_sumup.F90:46-53.
//
// They touch few beads (or, for Rg, need one reduction per replica), so they
// run as two small kernels of the per-phase form: per-replica statistics
// (centre of mass, Rg), then per bead the sum of its own terms, added to the
// bonded-force buffer before the pair pass adds the rest.  Each bead owns
// a list of the term ends it carries (built on the host), so there are no
// atomics.
#pragma once

namespace sis {

// one end of a term acting on bead `bead`
enum : int { XT_DCF1 = 0, XT_DCF2 = 1, XT_FR1 = 2, XT_FR2 = 3, XT_REST_I = 4, XT_REST_J = 5 };

struct ExtDev {
  int stage, rg, ndcf, nfr, nrest;
  double stage_sigma, stage_eps, rg_k;
  const double *rg0;          // [R]
  const int *term_off;        // [N+1] per-bead slices of terms
  const int2 *terms;          // (kind, index) per term end
  const int *dcf_pairs;       // [2*ndcf] 0-based beads
  const double *dcf_f;        // [3*ndcf*R]
  const int *fr_pairs;        // [2*nfr]
  const double *fr_k;         // [2*nfr]
  const double *fr_init;      // [6*nfr] (3, 2, nfr)
  const double *fr_velo;      // [6*nfr]
  const int *rest_id;         // [3*nrest] 0-based beads, type
  const double *rest_rcut;    // [nrest]
  const double *rest_para;    // [3*nrest]
  double *rg_stat;            // [4*R] centre of mass, Rg
  double *ene;                // [4*R] stage, tweezers, Rg, restraint
  int *err;                   // [1] set when a bead is below 0.85 stage_sigma in energy_stage
};

// centre of mass and Rg of each replica (src/force_rg.F90:17-27)
template <int B>
__global__ void __launch_bounds__(B) k_rg_stats(Params P, Ptrs D, ExtDev X) {
  __shared__ double s[3][B];
  const int r = blockIdx.x, N = P.N;
  double cx = 0.0, cy = 0.0, cz = 0.0;
  for (int a = threadIdx.x; a < N; a += B) {
    const double4 p = D.pos[(size_t)r * N + a];
    cx += p.x; cy += p.y; cz += p.z;
  }
  s[0][threadIdx.x] = cx; s[1][threadIdx.x] = cy; s[2][threadIdx.x] = cz;
  __syncthreads();
  for (int off = B / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 0; k < 3; k++) s[k][threadIdx.x] += s[k][threadIdx.x + off];
    __syncthreads();
  }
  const double mx = s[0][0] / N, my = s[1][0] / N, mz = s[2][0] / N;
  __syncthreads();
  double s2 = 0.0;
  for (int a = threadIdx.x; a < N; a += B) {
    const double4 p = D.pos[(size_t)r * N + a];
    s2 += dot_nc(p.x - mx, p.y - my, p.z - mz, p.x - mx, p.y - my, p.z - mz);
  }
  s[0][threadIdx.x] = s2;
  __syncthreads();
  for (int off = B / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[0][threadIdx.x] += s[0][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    X.rg_stat[4 * r] = mx; X.rg_stat[4 * r + 1] = my; X.rg_stat[4 * r + 2] = mz;
    X.rg_stat[4 * r + 3] = sqrt(s[0][0] / N);
  }
}

// sigmoid restraint (src/force_restraint.F90:18-40): v = x_i - x_j (minimum
// image), coef = eps/(2 s) (1 - tanh^2((r - d)/s)) / r inside r_cut
__device__ __forceinline__ bool rest_coef(const ExtDev &X, const Params &P, const Ptrs &D, int r, int k,
                                          double &vx, double &vy, double &vz, double &coef, double &rr) {
  const int i = X.rest_id[3 * k], j = X.rest_id[3 * k + 1];
  const double4 a = D.pos[(size_t)r * P.N + i], b = D.pos[(size_t)r * P.N + j];
  pbc_vec(D.box + 3 * r, P.pbc, a.x, a.y, a.z, b.x, b.y, b.z, vx, vy, vz);
  rr = sqrt(dot_nc(vx, vy, vz, vx, vy, vz));
  if (rr >= X.rest_rcut[k]) return false;
  const double eps = X.rest_para[3 * k], d = X.rest_para[3 * k + 1], s = X.rest_para[3 * k + 2];
  const double t = tanh((rr - d) / s);
  coef = 0.5 * eps / s * (1.0 - t * t) / rr;
  return true;
}

// per bead: the sum of its external terms, added to D.fbond
template <int B>
__global__ void __launch_bounds__(B) k_external(Params P, Ptrs D, ExtDev X, int64_t istep, const int *mask) {
  const int a = blockIdx.x * B + threadIdx.x, r = blockIdx.y, N = P.N;
  if (a >= N || (mask && !mask[r])) return;
  const size_t g = (size_t)r * N + a;
  const double4 p = D.pos[g];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (X.stage) {  // src/force_stage.F90:17-30
    double z = p.z;
    if (z < 2.220446049250313e-16) z = 2.220446049250313e-16;
    z = X.stage_sigma / z;
    const double z2 = z * z, z3 = z2 * z, z6 = z3 * z3, z7 = z6 * z, z13 = z6 * z7;
    fz += 24.0 * X.stage_eps / X.stage_sigma * (2.0 * z13 - z7);
  }
  for (int t = X.term_off[a]; t < X.term_off[a + 1]; t++) {
    const int2 e = X.terms[t];
    const int k = e.y;
    switch (e.x) {
      case XT_DCF1:
      case XT_DCF2: {  // src/force_tweezers.F90:30-36
        const double *f = X.dcf_f + 3 * ((size_t)r * X.ndcf + k);
        const double sg = e.x == XT_DCF1 ? 1.0 : -1.0;
        fx += sg * f[0]; fy += sg * f[1]; fz += sg * f[2];
        break;
      }
      case XT_FR1:
      case XT_FR2: {  // src/force_tweezers.F90:38-48
        const int end = e.x == XT_FR1 ? 0 : 1;
        const double *p0 = X.fr_init + 6 * k + 3 * end, *v0 = X.fr_velo + 6 * k + 3 * end;
        const double st = (double)istep;
        double vx, vy, vz;
        pbc_vec(D.box + 3 * r, P.pbc, p0[0] + v0[0] * st, p0[1] + v0[1] * st, p0[2] + v0[2] * st, p.x, p.y, p.z,
                vx, vy, vz);
        const double kk = X.fr_k[2 * k + end];
        fx += kk * vx; fy += kk * vy; fz += kk * vz;
        break;
      }
      default: {  // restraints
        double vx, vy, vz, coef, rr;
        if (!rest_coef(X, P, D, r, k, vx, vy, vz, coef, rr)) break;
        const double sg = e.x == XT_REST_I ? -1.0 : 1.0;
        fx += sg * coef * vx; fy += sg * coef * vy; fz += sg * coef * vz;
        break;
      }
    }
  }
  if (X.rg) {  // src/force_rg.F90:29-39
    const double *st = X.rg_stat + 4 * r;
    const double rg = st[3], r0 = X.rg0[r];
    if (!(X.rg == 2 && rg <= r0)) {
      const double factor = -X.rg_k * (rg - r0) / (rg * N);
      fx += factor * (p.x - st[0]); fy += factor * (p.y - st[1]); fz += factor * (p.z - st[2]);
    }
  }
  double4 fb = D.fbond[g];
  fb.x += fx; fb.y += fy; fb.z += fz;
  D.fbond[g] = fb;
}

// energies of the external terms per replica (src/energy_stage.F90,
// energy_tweezers.F90, energy_rg.F90, energy_restraint.F90); the Rg
// statistics come from k_rg_stats
template <int B>
__global__ void __launch_bounds__(B) k_external_energy(Params P, Ptrs D, ExtDev X) {
  __shared__ double s[B];
  const int r = blockIdx.x, N = P.N;
  double e = 0.0;
  if (X.stage) {
    for (int a = threadIdx.x; a < N; a += B) {
      const double z0 = D.pos[(size_t)r * N + a].z;
      if (z0 < 0.85 * X.stage_sigma) atomicOr(X.err, 1);
      const double z = X.stage_sigma / z0, z2 = z * z, z6 = z2 * z2 * z2;
      e += 4.0 * X.stage_eps * (z6 * z6 - z6);
    }
  }
  s[threadIdx.x] = e;
  __syncthreads();
  for (int off = B / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double etw = 0.0;
  for (int k = 0; k < X.ndcf; k++) {
    const int i = X.dcf_pairs[2 * k], j = X.dcf_pairs[2 * k + 1];
    const double4 a = D.pos[(size_t)r * N + i], b = D.pos[(size_t)r * N + j];
    double vx, vy, vz;
    pbc_vec(D.box + 3 * r, P.pbc, a.x, a.y, a.z, b.x, b.y, b.z, vx, vy, vz);
    const double *f = X.dcf_f + 3 * ((size_t)r * X.ndcf + k);
    etw -= dot_nc(vx, vy, vz, f[0], f[1], f[2]);
  }
  double erg = 0.0;
  if (X.rg) {
    const double rg = X.rg_stat[4 * r + 3], r0 = X.rg0[r];
    if (!(X.rg == 2 && rg <= r0)) erg = 0.5 * X.rg_k * (rg - r0) * (rg - r0);
  }
  double ers = 0.0;
  for (int k = 0; k < X.nrest; k++) {
    double vx, vy, vz, coef, rr;
    if (!rest_coef(X, P, D, r, k, vx, vy, vz, coef, rr)) continue;
    const double eps = X.rest_para[3 * k], d = X.rest_para[3 * k + 1], sg = X.rest_para[3 * k + 2];
    ers -= 0.5 * eps * (1.0 - tanh((rr - d) / sg));
  }
  X.ene[4 * r] = s[0];
  X.ene[4 * r + 1] = etw;
  X.ene[4 * r + 2] = erg;
  X.ene[4 * r + 3] = ers;
}

}  // namespace sis
