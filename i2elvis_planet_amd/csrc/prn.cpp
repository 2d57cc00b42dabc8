// prn.cpp -- the reference's binary .prn checkpoint for the host state:
// the byte layout of saver() (loadmart.c:671-809) and loader()
// (loadmart.c:391-540), so that files written here load in the reference and
// in its plotting script (support_files/plot2d.py:157-283), and the
// reference's files restart here.
//
// Layout (native little-endian; sizes int 4, long 8, float 4, double 8):
//   4 size bytes; xnumx ynumy mnumx mnumy marknum (long); xsize ysize
//   gamma_eff memory_fe memory_si por_init (double); growth_model (long);
//   gr_init (double); znumz corr2d3d (long); pinit pkf[4] GXKOEF GYKOEF
//   tmp_ambient (double); timeexit/yr, al2627*1e5, fe6056*1e8 (double);
//   rocknum (int); bondnum (long); n0 (int); timesum/yr (double);
//   rock table: markim (int) then 24 double arrays of rocknum;
//   per node (112 B): pr vx vy (float), bondm[3m..3m+2] (long), exx eyy exy
//   sxx syy sxy ro nu nd po mu ep et tk (float), bondm[3N+m] (long), cp kt ht
//   (float); gx, gy (float); per BC row 1..bondnum-1: bondv[4] (float),
//   bondn[3] (long); per marker 0..marknum -- one record more than there are
//   markers, as the reference's loops run to <= marknum -- (39 B): x y k v hi
//   por gr tmax acc (float), mg_old mg_time type (char).
// Quantities this build does not model are written as zero and skipped on
// reading: gamma_eff, memory_*, growth_model, gr_init, znumz, corr2d3d, pinit,
// pkf, timeexit, markim, marks0/1, markll, marka/b/e*, markht, nodes po, mu,
// ep, markers hi, gr, acc, mg_*.  The extra marker record is all zero.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/i2e.h"

namespace i2e {
void set_error(const char *fmt, ...);
}

namespace {

constexpr double YEAR = 3.15576e+7;  // seconds per year (loadmart.c:734-737)
constexpr int NDROCK = 24;           // double arrays of the rock table

struct Writer {
  FILE *f = nullptr;
  bool ok = true;
  void put(const void *p, size_t n) {
    if (ok && fwrite(p, 1, n, f) != n) ok = false;
  }
  template <class T>
  void v(T x) {
    put(&x, sizeof x);
  }
};

struct Reader {
  FILE *f = nullptr;
  bool ok = true;
  void get(void *p, size_t n) {
    if (ok && fread(p, 1, n, f) != n) ok = false;
  }
  template <class T>
  T v() {
    T x{};
    get(&x, sizeof x);
    return x;
  }
  void skip(long n) {
    if (ok && fseek(f, n, SEEK_CUR) != 0) ok = false;
  }
};

// rock-table array k (0..23, saver order) of type t as this build stores it
double rock_value(const i2e_rock &r, int k) {
  switch (k) {
    case 0: return r.n0;   // markn0
    case 1: return r.n1;   // markn1
    case 4: return r.nu0;  // marknu
    case 5: return r.dh;   // markdh
    case 6: return r.dv;   // markdv
    case 7: return r.ss;   // markss
    case 8: return r.mm;   // markmm
    case 16: return r.ro;  // markro
    case 17: return r.bb;  // markbb
    case 18: return r.aa;  // markaa
    case 19: return r.cp;  // markcp
    case 20: return r.kt;  // markkt
    case 21: return r.kf;  // markkf
    case 22: return r.kp;  // markkp
    default: return 0.0;   // marks0/1, markll, marka0..marke1, markht
  }
}
void set_rock_value(i2e_rock &r, int k, double x) {
  switch (k) {
    case 0: r.n0 = x; break;
    case 1: r.n1 = x; break;
    case 4: r.nu0 = x; break;
    case 5: r.dh = x; break;
    case 6: r.dv = x; break;
    case 7: r.ss = x; break;
    case 8: r.mm = x; break;
    case 16: r.ro = x; break;
    case 17: r.bb = x; break;
    case 18: r.aa = x; break;
    case 19: r.cp = x; break;
    case 20: r.kt = x; break;
    case 21: r.kf = x; break;
    case 22: r.kp = x; break;
    default: break;
  }
}

int read_head(Reader &r, i2e_prn_header *h) {
  unsigned char sz[4];
  r.get(sz, 4);
  if (!r.ok) return I2E_ERR_ARG;
  if (sz[0] != 4 || sz[1] != 8 || sz[2] != 4 || sz[3] != 8) {
    i2e::set_error("prn: variable sizes %d/%d/%d/%d in the file, expected 4/8/4/8 (int/long/float/double)", sz[0],
                   sz[1], sz[2], sz[3]);
    return I2E_ERR_UNSUPPORTED;
  }
  memset(h, 0, sizeof *h);
  h->xnumx = r.v<int64_t>();
  h->ynumy = r.v<int64_t>();
  h->mnumx = r.v<int64_t>();
  h->mnumy = r.v<int64_t>();
  h->marknum = r.v<int64_t>();
  h->xsize = r.v<double>();
  h->ysize = r.v<double>();
  r.skip(3 * 8);  // gamma_eff, memory_fe, memory_si
  h->por_init = r.v<double>();
  r.skip(8 + 8 + 8 + 8 + 8 + 4 * 8);  // growth_model, gr_init, znumz, corr2d3d, pinit, pkf
  h->gxkoef = r.v<double>();
  h->gykoef = r.v<double>();
  h->tmp_ambient = r.v<double>();
  r.skip(8);  // timeexit
  h->al2627_init = r.v<double>() * 1.0e-5;
  h->fe6056_init = r.v<double>() * 1.0e-8;
  h->rocknum = r.v<int32_t>();
  h->bondnum = r.v<int64_t>();
  h->n0 = r.v<int32_t>();
  h->timesum = r.v<double>() * YEAR;
  if (!r.ok) {
    i2e::set_error("prn: file ends inside the header");
    return I2E_ERR_ARG;
  }
  if (h->xnumx < 2 || h->ynumy < 2 || h->marknum < 0 || h->bondnum < 1 || h->rocknum < 0) {
    i2e::set_error("prn: header sizes out of range");
    return I2E_ERR_ARG;
  }
  return I2E_OK;
}

}  // namespace

extern "C" {

int i2e_prn_write(const char *path, const i2e_params *p, const i2e_host_state *s, int n0) {
  if (!path || !p || !s || s->marknum < 0 || s->bondnum < 1) return I2E_ERR_ARG;
  if ((int64_t)p->xnumx * p->ynumy <= 0) return I2E_ERR_ARG;
  Writer w;
  w.f = fopen(path, "wb");
  if (!w.f) {
    i2e::set_error("prn: cannot open %s for writing", path);
    return I2E_ERR_ARG;
  }
  const int64_t X = p->xnumx, Y = p->ynumy, N = X * Y, M = s->marknum;
  const unsigned char sz[4] = {4, 8, 4, 8};
  w.put(sz, 4);
  w.v<int64_t>(X);
  w.v<int64_t>(Y);
  w.v<int64_t>(p->mnumx);
  w.v<int64_t>(p->mnumy);
  w.v<int64_t>(M);
  w.v<double>(p->xsize);
  w.v<double>(p->ysize);
  for (int k = 0; k < 3; k++) w.v<double>(0.0);  // gamma_eff, memory_fe, memory_si
  w.v<double>(p->por_init);
  w.v<int64_t>(0);                                 // growth_model
  w.v<double>(0.0);                                // gr_init
  w.v<int64_t>(0);                                 // znumz
  w.v<int64_t>(0);                                 // corr2d3d
  for (int k = 0; k < 5; k++) w.v<double>(0.0);   // pinit, pkf[4]
  w.v<double>(s->gxkoef);
  w.v<double>(p->gykoef);
  w.v<double>(p->tmp_ambient);
  w.v<double>(0.0);  // timeexit
  w.v<double>(p->al2627_init * 1.0e+5);
  w.v<double>(p->fe6056_init * 1.0e+8);
  const int32_t rocknum = p->ntypes > 0 && p->ntypes <= I2E_MAXTYPES ? p->ntypes : I2E_MAXTYPES;
  w.v<int32_t>(rocknum);
  w.v<int64_t>(s->bondnum);
  w.v<int32_t>(n0);
  w.v<double>(s->timesum / YEAR);
  for (int t = 0; t < rocknum; t++) w.v<int32_t>(p->immobile[t] ? 1 : 0);  // markim
  for (int k = 0; k < NDROCK; k++)
    for (int t = 0; t < rocknum; t++) w.v<double>(rock_value(p->rock[t], k));
  for (int64_t m = 0; m < N && w.ok; m++) {
    w.v<float>((float)s->pr[m]);
    w.v<float>((float)s->vx[m]);
    w.v<float>((float)s->vy[m]);
    for (int c = 0; c < 3; c++) w.v<int64_t>(s->bondm[3 * m + c]);
    const double *f14[14] = {s->exx, s->eyy, s->exy, s->sxx, s->syy, s->sxy, s->ro,
                             s->nu,  s->nd,  nullptr, nullptr, nullptr, s->et, s->tk};  // po, mu, ep: 0
    for (int k = 0; k < 14; k++) w.v<float>(f14[k] ? (float)f14[k][m] : 0.0f);
    w.v<int64_t>(s->bondm[3 * N + m]);
    w.v<float>((float)s->cp[m]);
    w.v<float>((float)s->kt[m]);
    w.v<float>((float)s->ht[m]);
  }
  for (int64_t i = 0; i < X; i++) w.v<float>((float)s->gx[i]);
  for (int64_t j = 0; j < Y; j++) w.v<float>((float)s->gy[j]);
  for (int64_t b = 1; b < s->bondnum; b++) {
    for (int c = 0; c < 4; c++) w.v<float>((float)s->bondv[4 * b + c]);
    for (int c = 0; c < 3; c++) w.v<int64_t>(s->bondn[3 * b + c]);
  }
  for (int64_t i = 0; i <= M && w.ok; i++) {
    const bool in = i < M;
    const float rec[9] = {in ? s->markx[i] : 0.0f, in ? s->marky[i] : 0.0f, in ? s->markk[i] : 0.0f,
                          in ? s->markv[i] : 0.0f, 0.0f,                     in ? s->markpor[i] : 0.0f,
                          0.0f,                    in ? s->marktmax[i] : 0.0f, 0.0f};
    w.put(rec, sizeof rec);
    // magnetisation state and its time in Ma, each one byte as the saver
    // casts them (loadmart.c:805-806)
    int8_t mg = 0, mgt = 0;
    if (in && s->markmg_old) mg = s->markmg_old[i];
    if (in && s->markmg_time) {
      const double t = s->markmg_time[i];
      mgt = (int8_t)(t >= 127.0 ? 127 : t <= -128.0 ? -128 : (int)t);
    }
    const int8_t tail[3] = {mg, mgt, in ? s->markt[i] : (int8_t)0};
    w.put(tail, 3);
  }
  const bool ok = w.ok;
  if (fclose(w.f) != 0 || !ok) {
    i2e::set_error("prn: write to %s failed", path);
    return I2E_ERR_ARG;
  }
  return I2E_OK;
}

int i2e_prn_read_header(const char *path, i2e_prn_header *h) {
  if (!path || !h) return I2E_ERR_ARG;
  Reader r;
  r.f = fopen(path, "rb");
  if (!r.f) {
    i2e::set_error("prn: cannot open %s", path);
    return I2E_ERR_ARG;
  }
  const int st = read_head(r, h);
  fclose(r.f);
  return st;
}

int i2e_prn_read(const char *path, i2e_params *p, i2e_host_state *s, int *n0) {
  if (!path || !p || !s) return I2E_ERR_ARG;
  Reader r;
  r.f = fopen(path, "rb");
  if (!r.f) {
    i2e::set_error("prn: cannot open %s", path);
    return I2E_ERR_ARG;
  }
  i2e_prn_header h;
  int st = read_head(r, &h);
  if (st) {
    fclose(r.f);
    return st;
  }
  if (h.marknum != s->marknum || h.bondnum != s->bondnum || h.rocknum > I2E_MAXTYPES) {
    fclose(r.f);
    i2e::set_error("prn: file holds %lld markers / %lld BC rows / %d rock types; the state has room for %lld / %lld "
                   "/ %d", (long long)h.marknum, (long long)h.bondnum, h.rocknum, (long long)s->marknum,
                   (long long)s->bondnum, I2E_MAXTYPES);
    return I2E_ERR_CAPACITY;
  }
  const int64_t X = h.xnumx, Y = h.ynumy, N = X * Y, M = h.marknum;
  p->xnumx = X;
  p->ynumy = Y;
  p->mnumx = h.mnumx;
  p->mnumy = h.mnumy;
  p->xsize = h.xsize;
  p->ysize = h.ysize;
  p->por_init = h.por_init;
  p->gykoef = h.gykoef;
  p->tmp_ambient = h.tmp_ambient;
  p->al2627_init = h.al2627_init;
  p->fe6056_init = h.fe6056_init;
  if (h.rocknum > 0) p->ntypes = h.rocknum;
  s->gxkoef = h.gxkoef;
  s->timesum = h.timesum;
  if (n0) *n0 = h.n0;
  // markim: immobile rock types are never advected and are left out of the
  // temperature sums (markmart.c:27, 905)
  for (int t = 0; t < I2E_MAXTYPES; t++) p->immobile[t] = 0;
  for (int t = 0; t < h.rocknum; t++) {
    const int32_t im = r.v<int32_t>();
    if (t < I2E_MAXTYPES) p->immobile[t] = im != 0;
  }
  for (int k = 0; k < NDROCK; k++)
    for (int t = 0; t < h.rocknum; t++) set_rock_value(p->rock[t], k, r.v<double>());
  for (int64_t m = 0; m < N && r.ok; m++) {
    s->pr[m] = r.v<float>();
    s->vx[m] = r.v<float>();
    s->vy[m] = r.v<float>();
    for (int c = 0; c < 3; c++) s->bondm[3 * m + c] = r.v<int64_t>();
    double *f14[14] = {s->exx, s->eyy, s->exy, s->sxx, s->syy, s->sxy, s->ro,
                       s->nu,  s->nd,  nullptr, nullptr, nullptr, s->et, s->tk};
    for (int k = 0; k < 14; k++) {
      const float v = r.v<float>();
      if (f14[k]) f14[k][m] = v;
    }
    s->bondm[3 * N + m] = r.v<int64_t>();
    s->cp[m] = r.v<float>();
    s->kt[m] = r.v<float>();
    s->ht[m] = r.v<float>();
  }
  for (int64_t i = 0; i < X; i++) s->gx[i] = r.v<float>();
  for (int64_t j = 0; j < Y; j++) s->gy[j] = r.v<float>();
  for (int c = 0; c < 4; c++) s->bondv[c] = 0.0;
  for (int c = 0; c < 3; c++) s->bondn[c] = 0;
  for (int64_t b = 1; b < h.bondnum; b++) {
    for (int c = 0; c < 4; c++) s->bondv[4 * b + c] = r.v<float>();
    for (int c = 0; c < 3; c++) s->bondn[3 * b + c] = r.v<int64_t>();
  }
  for (int64_t i = 0; i < M && r.ok; i++) {
    float rec[9];
    r.get(rec, sizeof rec);
    int8_t tail[3];
    r.get(tail, 3);
    s->markx[i] = rec[0];
    s->marky[i] = rec[1];
    s->markk[i] = rec[2];
    s->markv[i] = rec[3];
    s->markpor[i] = rec[5];
    s->marktmax[i] = rec[7];
    s->markt[i] = tail[2];
    s->markp[i] = 0.0f;
    // the bytes as written (the reference's binary loader assigns the
    // accretion time it just read to both, loadmart.c:533-534)
    if (s->markmg_old) s->markmg_old[i] = tail[0];
    if (s->markmg_time) s->markmg_time[i] = tail[1];
  }
  r.skip(39);  // the record one past the last marker
  const bool ok = r.ok;
  fclose(r.f);
  if (!ok) {
    i2e::set_error("prn: %s ends early", path);
    return I2E_ERR_ARG;
  }
  // the reference keeps gp only in memory; the gravity solve starts from zero
  for (int64_t m = 0; m < N; m++) s->gp[m] = s->ga[m] = 0.0;
  return I2E_OK;
}

}  // extern "C"
