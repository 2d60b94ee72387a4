// Trajectory output: DCD files in the reference's layout, fed from the
// device without stalling the step loop.  Included by sis_gpu.hip after the
// context and the step drivers.
//
// File layout (src/dcd.F90:287-355 write_header, :538-596 write_onestep;
// the file is opened form='unformatted', access='stream', :115-116, so the
// block sizes are written explicitly and no record markers are added):
//   int32 84, "CORD", 20 int32 control words (all 0 but word 11 = 1 when
//   unit-cell blocks follow and word 20 = 24), int32 84;
//   int32 84, int32 1 (titles), 80 blanks, int32 84;
//   int32 4, int32 natom, int32 4;
//   per frame: [int32 48, six float64 (box1, 0, box2, 0, 0, box3), int32 48]
//   when the box is present, then for x, y and z in turn: int32 4N, N
//   float32 (coordinate minus the centre of geometry when fix_com_origin),
//   int32 4N.
//
// Device feed: at a save step the positions of all replicas are copied
// into one of two pinned host buffers on the engine's stream; a writer
// thread waits for the copy and writes one frame per replica file, so the
// next segment of steps runs while the frames are written.
#pragma once

#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

namespace {

void dcd_put(FILE *f, const void *p, size_t n) {
  if (fwrite(p, 1, n, f) != n) throw ApiError{SIS_ERR_ARG, "DCD write failed"};
}
void dcd_put_i32(FILE *f, int32_t v) { dcd_put(f, &v, 4); }

void dcd_header(FILE *f, int32_t natom, bool unitcell) {
  int32_t control[20] = {0};
  control[10] = unitcell ? 1 : 0;
  control[19] = 24;
  dcd_put_i32(f, 84);
  dcd_put(f, "CORD", 4);
  dcd_put(f, control, sizeof control);
  dcd_put_i32(f, 84);
  char title[80];
  memset(title, ' ', sizeof title);
  dcd_put_i32(f, 4 + 80);
  dcd_put_i32(f, 1);
  dcd_put(f, title, sizeof title);
  dcd_put_i32(f, 4 + 80);
  dcd_put_i32(f, 4);
  dcd_put_i32(f, natom);
  dcd_put_i32(f, 4);
  fflush(f);
}

// One frame; xyz(k, a) at xyz[stride * a + k] (stride 3: the reference's
// xyz(3, nmp); stride 4: the device's double4 layout).
void dcd_frame(FILE *f, int nmp, const double *xyz, int stride, bool unitcell, const double box[3], int fix_com,
               std::vector<float> &tmp) {
  if (unitcell) {
    const double cell[6] = {box[0], 0.0, box[1], 0.0, 0.0, box[2]};
    dcd_put_i32(f, 48);
    dcd_put(f, cell, sizeof cell);
    dcd_put_i32(f, 48);
  }
  double com[3] = {0.0, 0.0, 0.0};
  if (fix_com > 0) {
    for (int k = 0; k < 3; k++) {
      double s = 0.0;
      for (int a = 0; a < nmp; a++) s += xyz[(size_t)stride * a + k];
      com[k] = s / (double)nmp;
    }
  }
  tmp.resize(nmp);
  const int32_t num = nmp * 4;
  for (int k = 0; k < 3; k++) {
    for (int a = 0; a < nmp; a++) tmp[a] = (float)(xyz[(size_t)stride * a + k] - com[k]);
    dcd_put_i32(f, num);
    dcd_put(f, tmp.data(), (size_t)nmp * 4);
    dcd_put_i32(f, num);
  }
  fflush(f);
}

}  // namespace

struct DcdOut {
  std::vector<FILE *> f;
  int fix_com = 0;
  int64_t nstep_save = 0;
  bool unitcell = false;
  double box[3] = {0, 0, 0};      // current unit cell (follows box resizes)
  double sbox[2][3] = {};          // unit cell of the frame in each staging buffer
  int device = 0;
  // two pinned staging buffers of R*N double4
  double4 *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool busy[2] = {false, false};
  int next = 0;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> q;
  bool stop = false;
  std::string err;
  int64_t frames = 0;
};

namespace {

void dcd_writer_loop(sis_gpu_ctx *c, DcdOut *o) {
  (void)hipSetDevice(o->device);
  std::vector<float> tmp;
  for (;;) {
    int b;
    {
      std::unique_lock<std::mutex> lk(o->mu);
      o->cv.wait(lk, [o] { return o->stop || !o->q.empty(); });
      if (o->q.empty()) return;
      b = o->q.front();
      o->q.pop_front();
    }
    std::string e;
    if (hipEventSynchronize(o->ev[b]) != hipSuccess) e = "DCD staging copy failed";
    try {
      if (e.empty())
        for (int r = 0; r < c->R; r++)
          dcd_frame(o->f[r], c->N, reinterpret_cast<const double *>(o->buf[b] + (size_t)r * c->N), 4, o->unitcell,
                    o->sbox[b], o->fix_com, tmp);
    } catch (const ApiError &ex) {
      e = ex.msg;
    }
    std::lock_guard<std::mutex> lk(o->mu);
    if (!e.empty() && o->err.empty()) o->err = e;
    o->busy[b] = false;
    o->frames++;
    o->cv.notify_all();
  }
}

// queue the current positions of every replica as one frame each
void dcd_enqueue(sis_gpu_ctx *c) {
  DcdOut *o = c->dcd;
  int b;
  {
    std::unique_lock<std::mutex> lk(o->mu);
    b = o->next;
    o->cv.wait(lk, [o, b] { return !o->busy[b]; });
    if (!o->err.empty()) throw ApiError{SIS_ERR_ARG, o->err};
    o->busy[b] = true;
    o->next ^= 1;
    for (int k = 0; k < 3; k++) o->sbox[b][k] = o->box[k];
  }
  const size_t RN = (size_t)c->R * c->N;
  HIPCHK(hipMemcpyAsync(o->buf[b], c->pos_d.p, RN * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipEventRecord(o->ev[b], c->stream));
  std::lock_guard<std::mutex> lk(o->mu);
  o->q.push_back(b);
  o->cv.notify_all();
}

// later frames carry the context's current box (fdcd%box after a resize,
// src/job_md.F90:582)
void dcd_set_box(sis_gpu_ctx *c) {
  if (!c->dcd) return;
  std::lock_guard<std::mutex> lk(c->dcd->mu);
  for (int k = 0; k < 3; k++) c->dcd->box[k] = c->prm.box[k];
}

// wait for queued frames; rethrows a writer failure
void dcd_drain(DcdOut *o) {
  std::unique_lock<std::mutex> lk(o->mu);
  o->cv.wait(lk, [o] { return o->q.empty() && !o->busy[0] && !o->busy[1]; });
  if (!o->err.empty()) {
    const std::string e = o->err;
    o->err.clear();
    throw ApiError{SIS_ERR_ARG, e};
  }
}

void dcd_shutdown(sis_gpu_ctx *c) {
  DcdOut *o = c->dcd;
  if (!o) return;
  {
    std::lock_guard<std::mutex> lk(o->mu);
    o->stop = true;
    o->cv.notify_all();
  }
  if (o->th.joinable()) o->th.join();
  for (FILE *f : o->f)
    if (f) fclose(f);
  for (int b = 0; b < 2; b++) {
    if (o->buf[b]) (void)hipHostFree(o->buf[b]);
    if (o->ev[b]) (void)hipEventDestroy(o->ev[b]);
  }
  delete o;
  c->dcd = nullptr;
}

}  // namespace

// Base-pair output (src/write_bp.F90:23-51 and :53-82; header written at
// open by src/init_out_files.F90:102-104, stream access): per save step,
// records (int16 i, int16 j, float32 E) of the formed pairs in list order,
// those with E <= -kT for .bp (kind 0) or all for .bpall (kind 1), then a
// (0, 0, 0.0) terminator.
struct BpOut {
  std::vector<FILE *> f;
  int kind = 0;
  int64_t nstep_save = 0;
};

namespace {

void bp_header(FILE *f) {
  dcd_put_i32(f, 2);
  dcd_put_i32(f, 4);
  fflush(f);
}

void bp_records(FILE *f, int64_t n, const int32_t *ij, int ij_stride, const uint8_t *status, const double *ene,
                double kT, int kind) {
  for (int64_t k = 0; k < n; k++) {
    if (!status[k]) continue;
    if (kind == 0 && !(ene[k] <= -kT)) continue;
    const int16_t i = (int16_t)ij[(size_t)ij_stride * k], j = (int16_t)ij[(size_t)ij_stride * k + 1];
    const float e = (float)ene[k];
    dcd_put(f, &i, 2);
    dcd_put(f, &j, 2);
    dcd_put(f, &e, 4);
  }
  const int16_t z = 0;
  const float ez = 0.0f;
  dcd_put(f, &z, 2);
  dcd_put(f, &z, 2);
  dcd_put(f, &ez, 4);
  fflush(f);
}

// one save step of every replica, from the packed candidate arrays
void bp_write_now(sis_gpu_ctx *c) {
  BpOut *o = c->bpo;
  const size_t nb = c->nbp_total;
  std::vector<int32_t> mp(nb * 8);
  std::vector<uint8_t> st(nb);
  std::vector<double> en(nb);
  if (nb) {
    HIPCHK(hipMemcpyAsync(mp.data(), c->bp_mp_d.p, nb * 8 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(st.data(), c->bp_status_d.p, nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(en.data(), c->ene_bp_d.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  // candidates of replica r: [off[r N], off[(r + 1) N]) of the per-bead offsets
  const std::vector<int> ro = packed_rep_off(c);
  for (int r = 0; r < c->R; r++) {
    const int64_t b0 = ro[r], b1 = ro[r + 1];
    bp_records(o->f[r], b1 - b0, mp.data() + (size_t)b0 * 8, 8, st.data() + b0, en.data() + b0,
               c->tK[r] * KB_KCAL, o->kind);
  }
}

void bp_shutdown(sis_gpu_ctx *c) {
  if (!c->bpo) return;
  for (FILE *f : c->bpo->f)
    if (f) fclose(f);
  delete c->bpo;
  c->bpo = nullptr;
}

}  // namespace

extern "C" {

int sis_bp_write_file(const char *path, int64_t n, const int32_t *ij, const uint8_t *status, const double *ene,
                      int32_t nsets, double tempK, int32_t kind) {
  API_BEGIN
  if (!path || n < 0 || nsets < 0 || (n > 0 && (!ij || !status || !ene)) || (kind != 0 && kind != 1))
    return fail(SIS_ERR_ARG, "invalid argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(SIS_ERR_ARG, std::string("cannot open ") + path);
  try {
    bp_header(f);
    for (int s = 0; s < nsets; s++)
      bp_records(f, n, ij + (size_t)2 * n * s, 2, status + (size_t)n * s, ene + (size_t)n * s, tempK * KB_KCAL, kind);
  } catch (...) {
    fclose(f);
    throw;
  }
  fclose(f);
  return SIS_OK;
  API_END
}

int sis_gpu_bp_open(sis_gpu_ctx *c, const char *const *paths, int32_t kind, int64_t nstep_save) {
  API_BEGIN
  if (!c || !paths || nstep_save < 0 || (kind != 0 && kind != 1)) return fail(SIS_ERR_ARG, "invalid argument");
  if (c->bpo) return fail(SIS_ERR_STATE, "base-pair output already open");
  c->bpo = new BpOut();
  try {
    c->bpo->kind = kind;
    c->bpo->nstep_save = nstep_save;
    for (int r = 0; r < c->R; r++) {
      if (!paths[r]) throw ApiError{SIS_ERR_ARG, "null path"};
      FILE *f = fopen(paths[r], "wb");
      if (!f) throw ApiError{SIS_ERR_ARG, std::string("cannot open ") + paths[r]};
      c->bpo->f.push_back(f);
      bp_header(f);
    }
  } catch (...) {
    bp_shutdown(c);
    throw;
  }
  return SIS_OK;
  API_END
}

int sis_gpu_bp_write(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!c->bpo) return fail(SIS_ERR_STATE, "base-pair output not open");
  bp_write_now(c);
  return SIS_OK;
  API_END
}

int sis_gpu_bp_close(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  bp_shutdown(c);
  return SIS_OK;
  API_END
}

int sis_dcd_write_file(const char *path, int32_t nmp, int32_t nframes, const double *xyz, int32_t unitcell,
                       const double *box, int32_t fix_com_origin) {
  API_BEGIN
  if (!path || nmp <= 0 || nframes < 0 || (nframes > 0 && !xyz) || (unitcell && !box))
    return fail(SIS_ERR_ARG, "invalid argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(SIS_ERR_ARG, std::string("cannot open ") + path);
  try {
    dcd_header(f, nmp, unitcell != 0);
    std::vector<float> tmp;
    for (int k = 0; k < nframes; k++)
      dcd_frame(f, nmp, xyz + (size_t)3 * nmp * k, 3, unitcell != 0, box, fix_com_origin, tmp);
  } catch (...) {
    fclose(f);
    throw;
  }
  fclose(f);
  return SIS_OK;
  API_END
}

int sis_gpu_dcd_open(sis_gpu_ctx *c, const char *const *paths, int32_t fix_com_origin, int64_t nstep_save) {
  API_BEGIN
  if (!c || !paths || nstep_save < 0) return fail(SIS_ERR_ARG, "invalid argument");
  if (c->dcd) return fail(SIS_ERR_STATE, "trajectory output already open");
  DcdOut *o = new DcdOut();
  c->dcd = o;
  try {
    o->fix_com = fix_com_origin;
    o->nstep_save = nstep_save;
    o->unitcell = c->prm.pbc != 0;
    for (int k = 0; k < 3; k++) o->box[k] = c->prm.box[k];
    HIPCHK(hipGetDevice(&o->device));
    const size_t RN = (size_t)c->R * c->N;
    for (int b = 0; b < 2; b++) {
      HIPCHK(hipHostMalloc((void **)&o->buf[b], RN * sizeof(double4), hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&o->ev[b], hipEventDisableTiming));
    }
    for (int r = 0; r < c->R; r++) {
      if (!paths[r]) throw ApiError{SIS_ERR_ARG, "null DCD path"};
      FILE *f = fopen(paths[r], "wb");
      if (!f) throw ApiError{SIS_ERR_ARG, std::string("cannot open ") + paths[r]};
      o->f.push_back(f);
      dcd_header(f, c->N, o->unitcell);
    }
    o->th = std::thread(dcd_writer_loop, c, o);
  } catch (...) {
    dcd_shutdown(c);
    throw;
  }
  return SIS_OK;
  API_END
}

int sis_gpu_dcd_write(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!c->dcd) return fail(SIS_ERR_STATE, "trajectory output not open");
  dcd_enqueue(c);
  dcd_drain(c->dcd);
  return SIS_OK;
  API_END
}

int sis_gpu_dcd_close(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!c->dcd) return SIS_OK;
  std::string e;
  try {
    dcd_drain(c->dcd);
  } catch (const ApiError &ex) {
    e = ex.msg;
  }
  dcd_shutdown(c);
  if (!e.empty()) return fail(SIS_ERR_ARG, e);
  return SIS_OK;
  API_END
}

}  // extern "C"

// ---------------------------------------------------------------------
// Energy lines (.out): write_out_header and write_out of
// src/write_out.F90:1-116.  The line format is
//   (i12, 1x, [i4, 1x,] f6.2, 7(1x, g13.6)) [+ (1x, g13.6) for Eele]
// with Ekin and energies(0:5) in the seven fields; the Fortran G, F and I
// edit descriptors are reproduced below (round to nearest).
struct OutFile {
  std::vector<FILE *> f;
  int64_t nstep_save = 0;
  bool replica = false, ele = false;
};

namespace {

std::string pad_left(const std::string &s, int w) {
  return (int)s.size() >= w ? s : std::string(w - s.size(), ' ') + s;
}

// Gw.d output editing (no exponent width): F editing with 4 trailing
// blanks when 0.1 <= |x| < 10^d after rounding to d significant digits,
// else Ew.d (0.dddddd E+ee, or 0.dddddd+eee beyond two exponent digits).
std::string fortran_g(double x, int w, int d) {
  char buf[96];
  if (std::isnan(x)) return pad_left("NaN", w);
  if (std::isinf(x)) return pad_left(x < 0 ? "-Infinity" : "Infinity", w);
  const bool neg = std::signbit(x);
  const double a = std::fabs(x);
  if (a == 0.0) {
    std::string f = std::string(neg ? "-" : "") + "0." + std::string(d - 1, '0');
    return pad_left(f, w - 4) + "    ";
  }
  snprintf(buf, sizeof buf, "%.*e", d - 1, a);
  const char *ep = strchr(buf, 'e');
  const int e10 = atoi(ep + 1);
  const int k = e10 + 1;   // 10^(k-1) <= |x| rounded < 10^k
  if (k >= 0 && k <= d) {
    char fb[96];
    snprintf(fb, sizeof fb, "%#*.*f", w - 4, d - k, x);
    return std::string(fb) + "    ";
  }
  std::string digits;
  for (const char *q = buf; q < ep; q++)
    if (*q != '.') digits += *q;
  const int ee = e10 + 1;
  char eb[16];
  if (std::abs(ee) <= 99) snprintf(eb, sizeof eb, "E%c%02d", ee < 0 ? '-' : '+', std::abs(ee));
  else snprintf(eb, sizeof eb, "%c%03d", ee < 0 ? '-' : '+', std::abs(ee));
  return pad_left(std::string(neg ? "-" : "") + "0." + digits + eb, w);
}

void out_header(FILE *f, bool replica, bool ele) {
  std::string h;
  int icol;
  if (replica) {
    h = "#(1)nframe   (2)R (3)T   (4)Ekin       (5)Epot       (6)Ebond     "
        " (7)Eangl      (8)Edih       (9)Ebp        (10)Eexv     ";
    icol = 10;
  } else {
    h = "#(1)nframe   (2)T   (3)Ekin       (4)Epot       (5)Ebond     "
        " (6)Eangl      (7)Edih       (8)Ebp        (9)Eexv      ";
    icol = 9;
  }
  if (ele) {
    char b[32];
    snprintf(b, sizeof b, " (%2d)Eele     ", icol + 1);
    h += b;
  }
  h += "\n";
  dcd_put(f, h.data(), h.size());
  fflush(f);
}

void out_line(FILE *f, int64_t istep, int label, double tK, double ekin, const double *e, bool replica, bool ele) {
  char b[64];
  std::string s;
  snprintf(b, sizeof b, "%12lld", (long long)istep);
  s += b;
  if (replica) {
    snprintf(b, sizeof b, " %4d", label);
    s += b;
  }
  snprintf(b, sizeof b, " %6.2f", tK);
  s += b;
  s += " " + fortran_g(ekin, 13, 6);
  for (int k = 0; k <= 5; k++) s += " " + fortran_g(e[k], 13, 6);
  if (ele) s += " " + fortran_g(e[6], 13, 6);
  s += "\n";
  dcd_put(f, s.data(), s.size());
  fflush(f);
}

void out_shutdown(sis_gpu_ctx *c) {
  if (!c->outf) return;
  for (FILE *f : c->outf->f)
    if (f) fclose(f);
  delete c->outf;
  c->outf = nullptr;
}

// energies (R x 16) and kinetic energies of every replica at the current
// step, labels and temperatures as they are now
void out_lines(sis_gpu_ctx *c, const double *e16, const double *ekin, const std::vector<int> &labels,
               const std::vector<double> &temps) {
  for (int r = 0; r < c->R; r++)
    out_line(c->outf->f[r], c->istep, labels[r], temps[r], ekin[r], e16 + (size_t)r * 16, c->outf->replica,
             c->outf->ele);
}

}  // namespace

extern "C" {

int sis_out_write_file(const char *path, int32_t nline, const int64_t *isteps, const int32_t *labels,
                       const double *temps, const double *ekin, const double *energies, int32_t replica,
                       int32_t ele) {
  API_BEGIN
  if (!path || nline < 0 || (nline > 0 && (!isteps || !temps || !ekin || !energies || (replica && !labels))))
    return fail(SIS_ERR_ARG, "invalid argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(SIS_ERR_ARG, std::string("cannot open ") + path);
  try {
    out_header(f, replica != 0, ele != 0);
    for (int k = 0; k < nline; k++)
      out_line(f, isteps[k], replica ? labels[k] : 0, temps[k], ekin[k], energies + (size_t)11 * k, replica != 0,
               ele != 0);
  } catch (...) {
    fclose(f);
    throw;
  }
  fclose(f);
  return SIS_OK;
  API_END
}

int sis_gpu_out_open(sis_gpu_ctx *c, const char *const *paths, int64_t nstep_save) {
  API_BEGIN
  if (!c || !paths || nstep_save < 0) return fail(SIS_ERR_ARG, "invalid argument");
  if (c->outf) return fail(SIS_ERR_STATE, "energy output already open");
  c->outf = new OutFile();
  try {
    c->outf->nstep_save = nstep_save;
    c->outf->replica = c->R > 1;
    c->outf->ele = c->P.ele_on != 0;
    for (int r = 0; r < c->R; r++) {
      if (!paths[r]) throw ApiError{SIS_ERR_ARG, "null path"};
      FILE *f = fopen(paths[r], "wb");
      if (!f) throw ApiError{SIS_ERR_ARG, std::string("cannot open ") + paths[r]};
      c->outf->f.push_back(f);
      out_header(f, c->outf->replica, c->outf->ele);
    }
  } catch (...) {
    out_shutdown(c);
    throw;
  }
  return SIS_OK;
  API_END
}

// Tweezers' force-ramp lines (.twz; src/init_out_files.F90:133-139,
// src/force_tweezers.F90:50-59): written from the positions at the start
// of every step whose count is a multiple of nstep_save, for every replica
// in turn, into one file.
struct TwzOut {
  FILE *f = nullptr;
  int64_t nstep_save = 0;
};

namespace {

void twz_lines(sis_gpu_ctx *c, int64_t istep) {
  const int R = c->R, N = c->N, nfr = c->X.nfr;
  if (!c->twz || !c->twz->f || nfr <= 0) return;
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<double4> h((size_t)R * N);
  HIPCHK(hipMemcpy(h.data(), c->pos_d.p, h.size() * sizeof(double4), hipMemcpyDeviceToHost));
  std::vector<int> pairs(2 * nfr);
  std::vector<double> kk(2 * nfr), init(6 * nfr), velo(6 * nfr);
  HIPCHK(hipMemcpy(pairs.data(), c->X.fr_pairs, pairs.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(kk.data(), c->X.fr_k, kk.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(init.data(), c->X.fr_init, init.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(velo.data(), c->X.fr_velo, velo.size() * sizeof(double), hipMemcpyDeviceToHost));
  const double pN = (1.0 / 4184.0) * 6.02214076e23 * 1.0e-22;   // JOUL2KCAL_MOL * 1e-22
  auto norm = [](const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
  auto pvec = [&](const double *a, const double *b, double *v) {   // pbc_vec_d (src/pbc.F90:140-165)
    for (int k = 0; k < 3; k++) {
      v[k] = a[k] - b[k];
      if (c->prm.pbc) {
        const double L = c->prm.box[k], hL = 0.5 * L;
        if (v[k] > hL) v[k] -= L; else if (v[k] < -hL) v[k] += L;
      }
    }
  };
  std::string s;
  char b[64];
  for (int r = 0; r < R; r++)
    for (int q = 0; q < nfr; q++) {
      const double4 x1 = h[(size_t)r * N + pairs[2 * q]], x2 = h[(size_t)r * N + pairs[2 * q + 1]];
      const double b1[3] = {x1.x, x1.y, x1.z}, b2[3] = {x2.x, x2.y, x2.z};
      double p1[3], p2[3], v1[3], v2[3], d12[3], dp[3];
      for (int k = 0; k < 3; k++) {
        p1[k] = init[6 * q + k] + velo[6 * q + k] * (double)istep;
        p2[k] = init[6 * q + 3 + k] + velo[6 * q + 3 + k] * (double)istep;
        d12[k] = b2[k] - b1[k];
        dp[k] = p2[k] - p1[k];
      }
      pvec(p1, b1, v1);
      pvec(p2, b2, v2);
      snprintf(b, sizeof b, "%13lld %3d", (long long)istep, q + 1);
      s += b;
      const double vals[6] = {norm(dp), norm(d12), norm(v1), norm(v2), kk[2 * q] * norm(v1) / pN,
                              kk[2 * q + 1] * norm(v2) / pN};
      for (double v : vals) s += " " + fortran_g(v, 15, 8);
      s += "\n";
    }
  if (fwrite(s.data(), 1, s.size(), c->twz->f) != s.size()) throw ApiError{SIS_ERR_ARG, "write failed (.twz)"};
}

void twz_shutdown(sis_gpu_ctx *c) {
  if (!c->twz) return;
  if (c->twz->f) fclose(c->twz->f);
  delete c->twz;
  c->twz = nullptr;
}

}  // namespace

int sis_gpu_twz_open(sis_gpu_ctx *c, const char *path, int64_t nstep_save) {
  API_BEGIN
  if (!c || !path || nstep_save <= 0) return fail(SIS_ERR_ARG, "invalid argument");
  if (c->twz) return fail(SIS_ERR_STATE, "tweezers output already open");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(SIS_ERR_ARG, std::string("cannot open ") + path);
  const char *hdr = "#  (1)step  (2)id (3)d[traps]   (4)d[beads]   (5)d[t1-b1]   (6)d[t2-b2]   (7)f[b1]      (8)f[b2]\n";
  fputs(hdr, f);
  c->twz = new TwzOut();
  c->twz->f = f;
  c->twz->nstep_save = nstep_save;
  return SIS_OK;
  API_END
}

int sis_gpu_twz_close(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  twz_shutdown(c);
  return SIS_OK;
  API_END
}

int sis_gpu_out_write(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!c->outf) return fail(SIS_ERR_STATE, "energy output not open");
  if (!c->lists_valid) return fail(SIS_ERR_STATE, "energy line written before rebuild");
  // the initial state (src/job_md.F90:280-291,331-333): energies at each
  // replica's temperature with a fresh limit MC, and Ekin
  const int R = c->R;
  std::vector<int> lab(R), all(R, 1);
  for (int r = 0; r < R; r++) lab[r] = c->rep2lab[r + 1];
  std::vector<double> ekin(R), e16((size_t)R * 16);
  kinetic_all(c, ekin.data());
  energy_phase(c, all, c->tK, c->P.nstep_bp_mc > 0 ? 1 : 0, e16.data());
  out_lines(c, e16.data(), ekin.data(), lab, c->tK);
  return SIS_OK;
  API_END
}

int sis_gpu_out_close(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  out_shutdown(c);
  return SIS_OK;
  API_END
}

}  // extern "C"

// ---------------------------------------------------------------------
// Reading a DCD file (read_header / read_onestep of src/dcd.F90:146-536)
// and the energy-only analysis of job_dcd (src/job_dcd.F90:59-80): every
// frame becomes the coordinates of replica 1, the lists are rebuilt, the
// energies evaluated (limit MC when nstep_bp_MC > 0), and a .out line
// written with frame number, label 1, T and Ekin = 0 (job_dcd computes no
// kinetic energy).
namespace {

struct DcdIn {
  FILE *f = nullptr;
  bool unitcell = false;
  int32_t natom = 0;
  ~DcdIn() {
    if (f) fclose(f);
  }
};

bool dcd_get(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }
int32_t dcd_get_i32(FILE *f) {
  int32_t v;
  if (!dcd_get(f, &v, 4)) throw ApiError{SIS_ERR_ARG, "truncated DCD file"};
  return v;
}
void dcd_skip_block(FILE *f) {
  const int32_t n = dcd_get_i32(f);
  if (n < 0 || fseek(f, n, SEEK_CUR) != 0 || dcd_get_i32(f) != n) throw ApiError{SIS_ERR_ARG, "bad DCD block"};
}

void dcd_open_read(DcdIn &d, const char *path) {
  d.f = fopen(path, "rb");
  if (!d.f) throw ApiError{SIS_ERR_ARG, std::string("cannot open ") + path};
  int32_t n = dcd_get_i32(d.f);
  char cord[4];
  int32_t control[20];
  if (n != 84 || !dcd_get(d.f, cord, 4) || memcmp(cord, "CORD", 4) != 0 || !dcd_get(d.f, control, sizeof control) ||
      dcd_get_i32(d.f) != 84)
    throw ApiError{SIS_ERR_ARG, "not a DCD file"};
  d.unitcell = control[10] == 1;
  dcd_skip_block(d.f);   // titles
  n = dcd_get_i32(d.f);
  if (n != 4) throw ApiError{SIS_ERR_ARG, "bad DCD atom block"};
  d.natom = dcd_get_i32(d.f);
  if (dcd_get_i32(d.f) != 4) throw ApiError{SIS_ERR_ARG, "bad DCD atom block"};
}

// next frame into xyz (natom x 3, double); false at the end of the file
bool dcd_read_frame(DcdIn &d, std::vector<double> &xyz, double box[3]) {
  int32_t n;
  if (!dcd_get(d.f, &n, 4)) return false;
  if (d.unitcell) {
    double cell[6];
    if (n != 48 || !dcd_get(d.f, cell, sizeof cell) || dcd_get_i32(d.f) != 48)
      throw ApiError{SIS_ERR_ARG, "bad DCD unit-cell block"};
    box[0] = cell[0];
    box[1] = cell[2];
    box[2] = cell[5];
    n = dcd_get_i32(d.f);
  }
  std::vector<float> c((size_t)d.natom);
  xyz.assign((size_t)3 * d.natom, 0.0);
  for (int k = 0; k < 3; k++) {
    if (k > 0) n = dcd_get_i32(d.f);
    if (n != 4 * d.natom || !dcd_get(d.f, c.data(), c.size() * 4) || dcd_get_i32(d.f) != n)
      throw ApiError{SIS_ERR_ARG, "bad DCD coordinate block"};
    for (int a = 0; a < d.natom; a++) xyz[(size_t)3 * a + k] = (double)c[a];
  }
  return true;
}

}  // namespace

extern "C" {

int sis_dcd_read_file(const char *path, int32_t nmp, int64_t max_frames, double *xyz, double *box,
                      int64_t *nframes) {
  API_BEGIN
  if (!path || !nframes || max_frames < 0 || (max_frames > 0 && !xyz)) return fail(SIS_ERR_ARG, "invalid argument");
  DcdIn d;
  dcd_open_read(d, path);
  if (d.natom != nmp) return fail(SIS_ERR_ARG, "atom count of the DCD file differs");
  std::vector<double> x;
  double b[3] = {-1.0, -1.0, -1.0};
  int64_t k = 0;
  while (dcd_read_frame(d, x, b)) {
    if (k < max_frames) std::copy(x.begin(), x.end(), xyz + (size_t)3 * nmp * k);
    k++;
  }
  if (box) std::copy(b, b + 3, box);
  *nframes = k;
  return SIS_OK;
  API_END
}

int sis_gpu_dcd_energies(sis_gpu_ctx *c, const char *path, const char *out_path, int64_t max_frames,
                         double *energies, int64_t *nframes) {
  API_BEGIN
  if (!c || !path || !nframes || max_frames < 0 || (max_frames > 0 && !energies))
    return fail(SIS_ERR_ARG, "invalid argument");
  DcdIn d;
  dcd_open_read(d, path);
  if (d.natom != c->N) return fail(SIS_ERR_ARG, "atom count of the DCD file differs from the system");
  FILE *fo = nullptr;
  if (out_path) {
    fo = fopen(out_path, "wb");
    if (!fo) return fail(SIS_ERR_ARG, std::string("cannot open ") + out_path);
  }
  try {
    if (fo) out_header(fo, false, c->P.ele_on != 0);
    const int N = c->N, R = c->R;
    std::vector<double> x;
    std::vector<double4> h((size_t)N);
    double b[3];
    std::vector<int> m(R, 0);
    m[0] = 1;
    std::vector<double> e16((size_t)R * 16);
    int64_t k = 0;
    while (dcd_read_frame(d, x, b)) {
      for (int a = 0; a < N; a++) h[a] = double4{x[(size_t)3 * a], x[(size_t)3 * a + 1], x[(size_t)3 * a + 2], 0.0};
      HIPCHK(hipMemcpyAsync(c->pos_d.p, h.data(), (size_t)N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
      // the DCD job's lists (src/main.F90:144-154)
      analysis_rebuild(c);
      energy_phase(c, m, c->tK, c->P.nstep_bp_mc > 0 ? 1 : 0, e16.data());
      if (k < max_frames) std::copy(e16.begin(), e16.begin() + 11, energies + (size_t)11 * k);
      k++;
      if (fo) out_line(fo, k, 1, c->tK[0], 0.0, e16.data(), false, c->P.ele_on != 0);
    }
    *nframes = k;
  } catch (...) {
    if (fo) fclose(fo);
    throw;
  }
  if (fo) fclose(fo);
  return SIS_OK;
  API_END
}

}  // extern "C"

// ---------------------------------------------------------------------
// job_check_force (src/job_check_force.F90:46-82) on the device: forces of
// replica irep against central differences of its total energy with step
// h along every coordinate (the lists stay valid: h is far below the
// skin), with the limit MC as the run is configured.  With out_path the
// per-bead lines of job_check_force are written.
extern "C" int sis_gpu_check_force(sis_gpu_ctx *c, int32_t irep, double h, const char *out_path, double *f_energy,
                                   double *f_force) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (irep < 1 || irep > c->R || !(h > 0.0)) return fail(SIS_ERR_ARG, "invalid argument");
  const int N = c->N;
  const int mc = c->P.nstep_bp_mc > 0 ? 1 : 0;
  // the CHECK_FORCE job's lists (src/main.F90:144-154); a later step
  // rebuilds the MD lists
  analysis_rebuild(c);
  std::vector<double> F((size_t)3 * N), fe((size_t)3 * N);
  if (sis_gpu_force(c, irep, mc, F.data()) != SIS_OK) return SIS_ERR_STATE;
  double4 *xd = c->pos_d.p + (size_t)(irep - 1) * N;
  std::vector<double4> x((size_t)N);
  HIPCHK(hipMemcpyAsync(x.data(), xd, (size_t)N * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double e[11];
  auto energy_at = [&](int a, int k, double v) {
    double4 p = x[a];
    (k == 0 ? p.x : (k == 1 ? p.y : p.z)) = v;
    HIPCHK(hipMemcpyAsync(xd + a, &p, sizeof p, hipMemcpyHostToDevice, c->stream));
    if (sis_gpu_energy(c, irep, -1.0, mc, e) != SIS_OK) throw ApiError{SIS_ERR_STATE, g_err};
    return e[0];
  };
  for (int a = 0; a < N; a++) {
    for (int k = 0; k < 3; k++) {
      const double x0 = k == 0 ? x[a].x : (k == 1 ? x[a].y : x[a].z);
      const double ep = energy_at(a, k, x0 + h);
      const double em = energy_at(a, k, x0 - h);
      fe[(size_t)3 * a + k] = -(ep - em) / (2.0 * h);
    }
    HIPCHK(hipMemcpyAsync(xd + a, &x[a], sizeof(double4), hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (out_path) {
    FILE *f = fopen(out_path, "wb");
    if (!f) return fail(SIS_ERR_ARG, std::string("cannot open ") + out_path);
    char b[64];
    for (int a = 0; a < N; a++) {
      std::string s;
      snprintf(b, sizeof b, "imp=%6d  f_energy", a + 1);
      s += b;
      for (int k = 0; k < 3; k++) s += " " + fortran_g(fe[(size_t)3 * a + k], 11, 4);
      s += "  f_force=";
      for (int k = 0; k < 3; k++) s += " " + fortran_g(F[(size_t)3 * a + k], 11, 4);
      snprintf(b, sizeof b, "\nf_energy - f_force%6d", a + 1);
      s += b;
      for (int k = 0; k < 3; k++) {
        snprintf(b, sizeof b, "%11.4f", fe[(size_t)3 * a + k] - F[(size_t)3 * a + k]);
        s += b;
      }
      s += "\n";
      if (fwrite(s.data(), 1, s.size(), f) != s.size()) {
        fclose(f);
        return fail(SIS_ERR_ARG, "write failed");
      }
    }
    fclose(f);
  }
  if (f_energy) std::copy(fe.begin(), fe.end(), f_energy);
  if (f_force) std::copy(F.begin(), F.end(), f_force);
  return SIS_OK;
  API_END
}

// ---------------------------------------------------------------------------
// Restart files (.rst): the blocks of src/write_rst.F90:25-145 in stream
// access, each an int32 block id, an int32 byte size and the block's fields
// (ids src/const_idx.F90:72-85).  Id 11 is this engine's own block: the
// counter-RNG seed (uint64) of the performance mode; the reference reader
// skips blocks it does not ask for (src/read_rst.F90:48-53).
// ---------------------------------------------------------------------------

struct RstOut {
  std::vector<std::string> paths;
  int64_t nstep_save = 0;
};

namespace {

enum : int32_t {
  RST_STEP = 1, RST_ANNEAL = 2, RST_XYZ = 3, RST_VELO = 4, RST_ACCEL = 5, RST_REPLICA = 6,
  RST_PRNG = 7, RST_PRNGREP = 8, RST_TWZ = 9, RST_PBC = 10, RST_SEED = 11
};
constexpr int32_t MTS_WORDS = 19 + 624;   // mt_stream's saved state (src/const.F90:40)

struct RstRec {
  int32_t grep = 1, nrep = 1, nmp = 0;
  int64_t istep = 0;
  const int32_t *rep2lab = nullptr;       // [nrep], written to replica 1's file when nrep > 1
  const double *xyz = nullptr, *vel = nullptr, *acc = nullptr;   // (3, nmp)
  const uint32_t *mts_rep = nullptr;      // [MTS_WORDS] or none
  const uint32_t *mts = nullptr;          // [MTS_WORDS] or none
  const double *box = nullptr;            // [3] or none (no PBC block)
  const uint64_t *seed = nullptr;         // or none
  int32_t ianneal = 0;                    // ANNEAL block (current schedule entry), 0 = none
};

struct ByteSink {
  std::vector<char> b;
  template <typename T> void put(T v) {
    const char *p = reinterpret_cast<const char *>(&v);
    b.insert(b.end(), p, p + sizeof v);
  }
  void raw(const void *p, size_t n) { b.insert(b.end(), (const char *)p, (const char *)p + n); }
};

std::vector<char> rst_bytes(const RstRec &r) {
  ByteSink s;
  const int32_t n3 = 3 * r.nmp;
  if (r.grep == 1 && r.nrep > 1) {
    s.put<int32_t>(RST_REPLICA);
    s.put<int32_t>(4 * (1 + 2 * r.nrep));
    s.put<int32_t>(r.nrep);
    for (int i = 1; i <= r.nrep; i++) { s.put<int32_t>(i); s.put<int32_t>(r.rep2lab[i - 1]); }
  }
  s.put<int32_t>(RST_STEP);
  s.put<int32_t>(4 + 8);
  s.put<int32_t>(1);          // istep_sim: reserved, always 1
  s.put<int64_t>(r.istep);
  if (r.ianneal > 0) {
    s.put<int32_t>(RST_ANNEAL);
    s.put<int32_t>(4);
    s.put<int32_t>(r.ianneal);
  }
  const double *arr[3] = {r.xyz, r.vel, r.acc};
  const int32_t ids[3] = {RST_XYZ, RST_VELO, RST_ACCEL};
  for (int k = 0; k < 3; k++) {
    s.put<int32_t>(ids[k]);
    s.put<int32_t>(4 * 2 + 8 * n3);
    s.put<int32_t>(r.grep);
    s.put<int32_t>(r.nmp);
    s.raw(arr[k], sizeof(double) * n3);
  }
  if (r.grep == 1 && r.nrep > 1 && r.mts_rep) {
    s.put<int32_t>(RST_PRNGREP);
    s.put<int32_t>(4 * MTS_WORDS);
    s.raw(r.mts_rep, 4 * MTS_WORDS);
  }
  if (r.mts) {
    s.put<int32_t>(RST_PRNG);
    s.put<int32_t>(4 + 4 * MTS_WORDS);
    s.put<int32_t>(r.grep);
    s.raw(r.mts, 4 * MTS_WORDS);
  }
  if (r.grep == 1 && r.box) {
    s.put<int32_t>(RST_PBC);
    s.put<int32_t>(8 * 3);
    s.raw(r.box, 8 * 3);
  }
  if (r.seed) {
    s.put<int32_t>(RST_SEED);
    s.put<int32_t>(8);
    s.put<uint64_t>(*r.seed);
  }
  return s.b;
}

void write_whole(const std::string &path, const std::vector<char> &b) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) throw ApiError{SIS_ERR_ARG, "cannot open " + path};
  const bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
  if (fclose(f) != 0 || !ok) throw ApiError{SIS_ERR_ARG, "write failed: " + path};
}

// Parsed blocks of one restart file.
struct RstIn {
  int32_t found = 0;      // bit (id - 1) set for every block seen
  int32_t grep = 0, nrep = 0, ianneal = 0;
  int64_t istep = 0;
  uint64_t seed = 0;
  std::vector<int32_t> rep2lab;
  std::vector<double> xyz, vel, acc;
  double box[3] = {0, 0, 0};
};

RstIn rst_parse(const std::string &path, int32_t nmp) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) throw ApiError{SIS_ERR_ARG, "cannot open " + path};
  std::vector<char> b;
  char tmp[65536];
  size_t got;
  while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) b.insert(b.end(), tmp, tmp + got);
  fclose(f);
  RstIn o;
  size_t p = 0;
  auto need = [&](size_t n) {
    if (p + n > b.size()) throw ApiError{SIS_ERR_ARG, "truncated restart file: " + path};
  };
  auto get32 = [&]() { need(4); int32_t v; std::memcpy(&v, b.data() + p, 4); p += 4; return v; };
  while (p < b.size()) {
    const int32_t id = get32(), size = get32();
    if (size < 0) throw ApiError{SIS_ERR_ARG, "bad block size in " + path};
    need((size_t)size);
    const size_t end = p + (size_t)size;
    // a block this reader knows must have its size; others (the MT state
    // blocks) are skipped
    auto bad = [&](const char *blk) { throw ApiError{SIS_ERR_ARG, std::string("bad ") + blk + " block in " + path}; };
    bool parsed = true;
    if (id == RST_STEP) {
      if (size != 12) bad("STEP");
      std::memcpy(&o.istep, b.data() + p + 4, 8);
    } else if (id == RST_XYZ || id == RST_VELO || id == RST_ACCEL) {
      if (size < 8) bad("XYZ/VELO/ACCEL");
      int32_t g, n;
      std::memcpy(&g, b.data() + p, 4);
      std::memcpy(&n, b.data() + p + 4, 4);
      if (n != nmp || (size_t)size != 8 + (size_t)24 * n)
        throw ApiError{SIS_ERR_ARG, "nmp is not consistent in the restart file " + path};
      o.grep = g;
      std::vector<double> &dst = id == RST_XYZ ? o.xyz : id == RST_VELO ? o.vel : o.acc;
      dst.resize((size_t)3 * n);
      std::memcpy(dst.data(), b.data() + p + 8, (size_t)24 * n);
    } else if (id == RST_REPLICA) {
      if (size < 4) bad("REPLICA");
      std::memcpy(&o.nrep, b.data() + p, 4);
      if (o.nrep < 1 || (size_t)size != 4 + (size_t)8 * o.nrep)
        throw ApiError{SIS_ERR_ARG, "bad REPLICA block in " + path};
      o.rep2lab.assign(o.nrep, 0);
      for (int k = 0; k < o.nrep; k++) {
        int32_t i, m;
        std::memcpy(&i, b.data() + p + 4 + 8 * k, 4);
        std::memcpy(&m, b.data() + p + 8 + 8 * k, 4);
        if (i < 1 || i > o.nrep) throw ApiError{SIS_ERR_ARG, "bad REPLICA block in " + path};
        o.rep2lab[i - 1] = m;
      }
    } else if (id == RST_ANNEAL) {
      if (size != 4) bad("ANNEAL");
      std::memcpy(&o.ianneal, b.data() + p, 4);
    } else if (id == RST_PBC) {
      if (size != 24) bad("PBC");
      std::memcpy(o.box, b.data() + p, 24);
    } else if (id == RST_SEED) {
      if (size != 8) bad("SEED");
      std::memcpy(&o.seed, b.data() + p, 8);
    } else {
      parsed = false;
    }
    if (parsed) o.found |= 1 << (id - 1);
    p = end;
  }
  return o;
}

// one file per replica from the device state (replica r's file holds grep r+1)
void rst_write_all(sis_gpu_ctx *c, const std::vector<std::string> &paths) {
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t RN = (size_t)c->R * c->N;
  std::vector<double> x(3 * RN), v(3 * RN), a(3 * RN);
  std::vector<double4> h(RN);
  auto down = [&](std::vector<double> &dst, double4 *src) {
    HIPCHK(hipMemcpy(h.data(), src, RN * sizeof(double4), hipMemcpyDeviceToHost));
    for (size_t g = 0; g < RN; g++) { dst[3 * g] = h[g].x; dst[3 * g + 1] = h[g].y; dst[3 * g + 2] = h[g].z; }
  };
  down(x, c->pos_d.p);
  down(v, c->vel_d.p);
  down(a, c->acc_d.p);
  std::vector<int32_t> lab(c->R);
  for (int r = 0; r < c->R; r++) lab[r] = c->rep2lab[r + 1];
  const uint64_t seed = c->opt.seed;
  for (int r = 0; r < c->R; r++) {
    RstRec q;
    q.grep = r + 1; q.nrep = c->R; q.nmp = c->N; q.istep = c->istep; q.rep2lab = lab.data();
    q.xyz = x.data() + (size_t)3 * c->N * r;
    q.vel = v.data() + (size_t)3 * c->N * r;
    q.acc = a.data() + (size_t)3 * c->N * r;
    q.box = c->prm.pbc ? c->prm.box : nullptr;
    q.seed = c->parity ? nullptr : &seed;   // the parity form's MT state belongs to the caller
    // written while a schedule entry remains (opt_anneal > 0, src/write_rst.F90:62-68)
    q.ianneal = (c->anneal_on && c->an_next > 0) ? c->ianneal : 0;
    write_whole(paths[r], rst_bytes(q));
  }
}

void rst_shutdown(sis_gpu_ctx *c) {
  delete c->rst;
  c->rst = nullptr;
}

}  // namespace

int sis_rst_write_file(const char *path, int32_t grep, int32_t nrep, int32_t nmp, int64_t istep,
                       const int32_t *rep2lab, const double *xyz, const double *velos, const double *accels,
                       const uint32_t *mts_rep, const uint32_t *mts, const double *box, const uint64_t *seed) {
  API_BEGIN
  if (!path || !xyz || !velos || !accels || nmp < 1 || nrep < 1 || grep < 1 || grep > nrep ||
      (grep == 1 && nrep > 1 && !rep2lab))
    return fail(SIS_ERR_ARG, "invalid argument");
  RstRec q;
  q.grep = grep; q.nrep = nrep; q.nmp = nmp; q.istep = istep; q.rep2lab = rep2lab;
  q.xyz = xyz; q.vel = velos; q.acc = accels; q.mts_rep = mts_rep; q.mts = mts; q.box = box; q.seed = seed;
  write_whole(path, rst_bytes(q));
  return SIS_OK;
  API_END
}

int sis_rst_read_file(const char *path, int32_t nmp, int32_t nrep, int32_t *blocks, int32_t *grep,
                      int64_t *istep, int32_t *rep2lab, double *xyz, double *velos, double *accels,
                      double *box, uint64_t *seed) {
  API_BEGIN
  if (!path || nmp < 1 || nrep < 1) return fail(SIS_ERR_ARG, "invalid argument");
  RstIn o = rst_parse(path, nmp);
  if ((o.found & (1 << (RST_REPLICA - 1))) && o.nrep != nrep)
    return fail(SIS_ERR_ARG, "nrep is not consistent in the restart file");
  if (blocks) *blocks = o.found;
  if (grep) *grep = o.grep;
  if (istep) *istep = o.istep;
  if (rep2lab && !o.rep2lab.empty()) std::copy(o.rep2lab.begin(), o.rep2lab.end(), rep2lab);
  if (xyz && !o.xyz.empty()) std::copy(o.xyz.begin(), o.xyz.end(), xyz);
  if (velos && !o.vel.empty()) std::copy(o.vel.begin(), o.vel.end(), velos);
  if (accels && !o.acc.empty()) std::copy(o.acc.begin(), o.acc.end(), accels);
  if (box) std::copy(o.box, o.box + 3, box);
  if (seed) *seed = o.seed;
  return SIS_OK;
  API_END
}

int sis_gpu_rst_open(sis_gpu_ctx *c, const char *const *paths, int64_t nstep_save_rst) {
  API_BEGIN
  if (!c || !paths || nstep_save_rst < 0) return fail(SIS_ERR_ARG, "invalid argument");
  if (c->rst) return fail(SIS_ERR_STATE, "restart output already open");
  RstOut *o = new RstOut();
  o->nstep_save = nstep_save_rst;
  for (int r = 0; r < c->R; r++) {
    if (!paths[r]) { delete o; return fail(SIS_ERR_ARG, "null path"); }
    o->paths.emplace_back(paths[r]);
  }
  c->rst = o;
  return SIS_OK;
  API_END
}

int sis_gpu_rst_write(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  if (!c->rst) return fail(SIS_ERR_STATE, "restart output not open");
  rst_write_all(c, c->rst->paths);
  return SIS_OK;
  API_END
}

int sis_gpu_rst_close(sis_gpu_ctx *c) {
  API_BEGIN
  if (!c) return fail(SIS_ERR_ARG, "null context");
  rst_shutdown(c);
  return SIS_OK;
  API_END
}

int sis_gpu_rst_read(sis_gpu_ctx *c, const char *const *paths) {
  API_BEGIN
  if (!c || !paths) return fail(SIS_ERR_ARG, "null argument");
  const int R = c->R, N = c->N;
  const size_t n3 = (size_t)3 * N;
  std::vector<double> x(n3 * R), v(n3 * R), a(n3 * R);
  std::vector<int> have(R, 0);
  int64_t istep = -1;
  double new_box[3] = {c->prm.box[0], c->prm.box[1], c->prm.box[2]};
  std::vector<int32_t> lab;
  int ianneal = -1;
  // every file is read and checked before the context changes
  for (int r = 0; r < R; r++) {
    if (!paths[r]) return fail(SIS_ERR_ARG, "null path");
    RstIn o = rst_parse(paths[r], N);
    const int need = (1 << (RST_XYZ - 1)) | (1 << (RST_VELO - 1)) | (1 << (RST_ACCEL - 1));
    if ((o.found & need) != need) return fail(SIS_ERR_ARG, std::string("XYZ, VELO or ACCEL block missing in ") + paths[r]);
    if (o.grep < 1 || o.grep > R || have[o.grep - 1]) return fail(SIS_ERR_ARG, "replica ids in the restart files are not 1..nrep");
    have[o.grep - 1] = 1;
    std::copy(o.xyz.begin(), o.xyz.end(), x.begin() + n3 * (o.grep - 1));
    std::copy(o.vel.begin(), o.vel.end(), v.begin() + n3 * (o.grep - 1));
    std::copy(o.acc.begin(), o.acc.end(), a.begin() + n3 * (o.grep - 1));
    if (o.found & (1 << (RST_STEP - 1))) istep = o.istep;
    if (o.found & (1 << (RST_ANNEAL - 1))) ianneal = o.ianneal;
    if (!o.rep2lab.empty()) {
      if (o.nrep != R) return fail(SIS_ERR_ARG, "nrep is not consistent in the restart file");
      lab = o.rep2lab;
    }
    if ((o.found & (1 << (RST_PBC - 1))) && c->prm.pbc) std::copy(o.box, o.box + 3, new_box);
  }
  if (istep < 0) return fail(SIS_ERR_ARG, "STEP block missing");
  if (R > 1 && lab.empty()) return fail(SIS_ERR_ARG, "REPLICA block missing");
  if (R > 1) {   // labels: a permutation of 1..nrep
    std::vector<int> seen(R + 1, 0);
    for (int r = 0; r < R; r++) {
      if (lab[r] < 1 || lab[r] > R || seen[lab[r]]) return fail(SIS_ERR_ARG, "bad labels in the REPLICA block");
      seen[lab[r]] = 1;
    }
  }
  if (!(new_box[0] > 0.0 && new_box[1] > 0.0 && new_box[2] > 0.0) && c->prm.pbc)
    return fail(SIS_ERR_ARG, "bad box in the PBC block");
  if (ianneal >= 0) c->rst_ianneal = ianneal;
  // the PBC block's box replaces the input box (init_pbc, src/pbc.F90:35-45)
  if (new_box[0] != c->prm.box[0] || new_box[1] != c->prm.box[1] || new_box[2] != c->prm.box[2]) {
    std::copy(new_box, new_box + 3, c->prm.box);
    dcd_set_box(c);
    upload_conditions(c);
    setup_cells(c);
    c->lists_valid = false;
  }
  int rc;
  if (R > 1 && (rc = sis_gpu_set_labels(c, lab.data())) != SIS_OK) return rc;
  if ((rc = sis_gpu_set_state(c, x.data(), v.data(), a.data())) != SIS_OK) return rc;
  if ((rc = sis_gpu_set_step(c, istep)) != SIS_OK) return rc;
  return SIS_OK;
  API_END
}
