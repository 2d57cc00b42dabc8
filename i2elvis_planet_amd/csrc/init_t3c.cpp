// init_t3c.cpp -- the initial state a reference run starts from, generated
// from an init.t3c file with the semantics of in2mart (in2mart.c:23-1453):
// grid lines, the marker lattice with its random jitter, the rock table,
// boundary-condition tables (regular boxes, irregular rectangles, circular
// sectors, grid-line definitions, cell marker sets), rock bodies on the
// markers and the initial temperature field, porosity and marker
// temperatures.  Host code only.
//
// The jitter uses the C library's rand() sequence without srand(), as the
// reference does: glibc's additive lagged-Fibonacci generator (TYPE_3,
// r[i] = r[i-3] + r[i-31], seeded by 16807^i mod 2^31-1 from seed 1, the
// first 310 values discarded, outputs r >> 1), written out here so that
// the markers land where the reference's in2mart puts them.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/i2e.h"

namespace i2e {
void set_error(const char *fmt, ...);
}

namespace {

struct GlibcRand {
  std::vector<uint32_t> ring;  // the generator's state, newest last
  GlibcRand() {
    int32_t s[34];
    s[0] = 1;
    for (int i = 1; i < 31; i++) {
      const int64_t v = (16807LL * s[i - 1]) % 2147483647LL;
      s[i] = (int32_t)(v < 0 ? v + 2147483647LL : v);
    }
    for (int i = 31; i < 34; i++) s[i] = s[i - 31];
    ring.assign(s, s + 34);
    for (int i = 34; i < 344; i++) ring.push_back(ring[i - 31] + ring[i - 3]);
  }
  int next() {
    const size_t i = ring.size();
    const uint32_t v = ring[i - 31] + ring[i - 3];
    ring.push_back(v);
    if (ring.size() > 4096) ring.erase(ring.begin(), ring.end() - 64);
    return (int)(v >> 1);
  }
};

struct Tokens {
  std::vector<std::string> t;
  size_t k = 0;
  bool load(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    std::string cur;
    int c;
    while ((c = fgetc(f)) != EOF) {
      if (c <= 32) {  // whitespace and control characters end a token (fscanf %s, then the trailing-<=32 cut)
        if (!cur.empty()) t.push_back(cur);
        cur.clear();
      } else {
        cur.push_back((char)c);
      }
    }
    if (!cur.empty()) t.push_back(cur);
    fclose(f);
    return true;
  }
  // the next token that does not start with '/' (ffscanf1, loadmart.c:878-903)
  bool next(std::string &s) {
    while (k < t.size() && t[k][0] == '/') k++;
    if (k >= t.size()) return false;
    s = t[k++];
    return true;
  }
};

struct Gen {
  // grid and run values from the header
  long X = 0, Y = 0, mnumx = 0, mnumy = 0, minmx = 0, maxmx = 0, minmy = 0, maxmy = 0;
  double xsize = 0, ysize = 0, por_init = 0, gr_init = 0, pinit = 0, gxkoef = 0, gykoef = 0, tmp_ambient = 0;
  double timeexit = 0, al2627 = 0, fe6056 = 0, timesum = 0;
  int nonstab = 0;
  double wamp = 0, wlen = 0, wymin = 0, wymid = 0, wymax = 0, xnonstab = 0, ynonstab = 0;
  double xstpx = 0, ystpy = 0;
  std::vector<double> gx, gy, tk, vx, vy;
  std::vector<float> mx, my, mk, mpor, mtmax;
  std::vector<signed char> mt;
  std::vector<int64_t> bondm, bondn;  // bondn: 3 per row
  std::vector<double> bondv;          // 4 per row
  int rocknum = 0;
  i2e_rock rock[I2E_MAXTYPES];
  uint8_t markim[I2E_MAXTYPES] = {};  // immobile rock types ('i' prefix)
  GlibcRand rnd;
  Tokens tk_in;
  std::string err;

  bool fail(const std::string &m) {
    if (err.empty()) err = m;
    return false;
  }
  bool tok(std::string &s) {
    if (!tk_in.next(s)) return fail("init: unexpected end of file");
    return true;
  }
  long ntok() {
    std::string s;
    return tok(s) ? atol(s.c_str()) : 0;
  }
  double dtok() {
    std::string s;
    return tok(s) ? atof(s.c_str()) : 0.0;
  }
  // "m<metres>" -> fraction of the box size, else a fraction already
  double frac(double size) {
    std::string s;
    if (!tok(s)) return 0;
    return s[0] == 'm' ? atof(s.c_str() + 1) / size : atof(s.c_str());
  }
  long nodes() const { return X * Y; }

  bool run(const char *path);
  bool header();
  bool markers();
  bool rocks();
  bool bcs();
  bool bodies();
  bool temperature();
  void finish();
  // a new BC row: {const, k1, k2, k3} and 1-based neighbours
  long add_row(double c) {
    const long b = (long)(bondv.size() / 4);
    bondv.insert(bondv.end(), {c, 0.0, 0.0, 0.0});
    bondn.insert(bondn.end(), {0, 0, 0});
    return b;
  }
};

// in2mart.c:66-127: the GRID_PARAMETERS block
bool Gen::header() {
  X = ntok();
  Y = ntok();
  mnumx = ntok();
  if (mnumx < 0) {
    mnumx = -mnumx;
    minmx = ntok();
    maxmx = ntok();
  }
  mnumy = ntok();
  if (mnumy < 0) {
    mnumy = -mnumy;
    minmy = ntok();
    maxmy = ntok();
  }
  xsize = dtok();
  ysize = dtok();
  dtok();  // gamma_eff
  dtok();  // memory_fe
  dtok();  // memory_si
  por_init = dtok();
  ntok();  // growth_model
  gr_init = dtok();
  ntok();  // znumz
  ntok();  // corr2d3d
  pinit = dtok();
  gxkoef = dtok();
  gykoef = dtok();
  tmp_ambient = dtok();
  timeexit = dtok() * 3.15576e+7;
  al2627 = dtok() * 1.0e-5;
  fe6056 = dtok() * 1.0e-8;
  timesum = dtok() * 3.15576e+7;
  nonstab = (int)ntok();
  if (nonstab < 0) {
    wamp = dtok();
    wlen = dtok();
    if (nonstab == -2) {
      wymin = dtok();
      wymid = dtok();
      wymax = dtok();
    }
  }
  if (nonstab > 0) {
    xnonstab = dtok() * xsize / ((double)(X - 1)) / ((double)(mnumx));
    ynonstab = dtok() * ysize / ((double)(Y - 1)) / ((double)(mnumy));
  }
  if (!err.empty()) return false;
  if (X < 3 || Y < 3 || mnumx < 1 || mnumy < 1 || !(xsize > 0) || !(ysize > 0))
    return fail("init: grid sizes out of range");
  xstpx = xsize / (double)(X - 1);
  ystpy = ysize / (double)(Y - 1);
  // grid lines, a running sum (in2mart.c:135-138)
  gx.assign(X, 0.0);
  gy.assign(Y, 0.0);
  for (long m1 = 1; m1 < X; m1++) gx[m1] = gx[m1 - 1] + xstpx;
  for (long m2 = 1; m2 < Y; m2++) gy[m2] = gy[m2 - 1] + ystpy;
  tk.assign(nodes(), 0.0);
  vx.assign(nodes(), 0.0);
  vy.assign(nodes(), 0.0);
  bondm.assign(4 * nodes(), 0);
  // row 0 means "no condition" in bondm; rows are numbered from 1
  // (bondnum = 1 before the first condition, in2mart.c:271)
  bondv.assign(4, 0.0);
  bondn.assign(3, 0);
  return true;
}

// in2mart.c:140-198: the marker lattice, its random jitter, the output name
bool Gen::markers() {
  std::string s;
  if (!tok(s)) return false;
  if (s[0] != '0') return fail("init: markers loaded from a file (" + s + ") are not supported");
  const double dx = xstpx / (double)(mnumx), dy = ystpy / (double)(mnumy);
  for (long m1 = minmx; m1 < X - 1 + maxmx; m1++)
    for (long m2 = minmy; m2 < Y - 1 + maxmy; m2++) {
      double x = (double)(m1)*xstpx + dx / 2.0;
      for (long m4 = 0; m4 < mnumx; m4++) {
        double y = (double)(m2)*ystpy + dy / 2.0;
        for (long m5 = 0; m5 < mnumy; m5++) {
          mx.push_back((float)(x));
          my.push_back((float)(y));
          y += dy;
        }
        x += dx;
      }
    }
  mt.assign(mx.size(), 0);
  if (nonstab > 0)
    for (size_t i = 0; i < mx.size(); i++)
      if (mt[i] < 100) {
        mx[i] += (float)(rnd.next() % (nonstab * 2 + 1) - nonstab) / ((float)(nonstab)) * (float)(xnonstab);
        my[i] += (float)(rnd.next() % (nonstab * 2 + 1) - nonstab) / ((float)(nonstab)) * (float)(ynonstab);
      }
  // output file name and type
  return tok(s) && tok(s);
}

// in2mart.c:208-246: the ROCKS table up to '~'
bool Gen::rocks() {
  std::string s;
  memset(rock, 0, sizeof rock);
  if (!tok(s)) return false;
  while (s[0] != '~') {
    const bool im = s[0] == 'i';
    const int n1 = atoi(s.c_str() + (im ? 1 : 0));
    if (n1 < 0 || n1 >= I2E_MAXTYPES) return fail("init: rock type " + s + " outside 0..63");
    markim[n1] = im ? 1 : 0;
    i2e_rock &r = rock[n1];
    r.n0 = dtok();
    r.n1 = dtok();
    dtok();  // marks0
    dtok();  // marks1
    r.nu0 = dtok();
    r.dh = dtok();
    r.dv = dtok();
    r.ss = dtok();
    r.mm = dtok();
    for (int k = 0; k < 7; k++) dtok();  // markll, marka0/1, markb0/1, marke0/1 (brittle, not modelled)
    r.ro = dtok();
    r.bb = dtok();
    r.aa = dtok();
    r.cp = dtok();
    r.kt = dtok();
    r.kf = dtok();
    r.kp = dtok();
    if (!tok(s)) return false;  // markht (per kg with a 'k' prefix): the heating is by isotopes here
    if (!tok(s)) return false;
    rocknum = rocknum > n1 + 1 ? rocknum : n1 + 1;
  }
  return err.empty();
}

// in2mart.c:267-984: boundary conditions up to '~'
bool Gen::bcs() {
  const double pival = 2.000 * asin(1.000);
  const long N = nodes(), N3 = 3 * N;
  std::string s;
  if (!tok(s)) return false;
  while (s[0] != '~') {
    int yn = 0;
    if (s[0] == 'R') yn = 1;
    if (s[0] == 'C') yn = 2;
    int n1 = -1;
    const char c0 = s[0], c1 = s.size() > 1 ? s[1] : 0;
    if (c1 == 'x') n1 = 0;
    if (c1 == 'y') n1 = 1;
    if (c0 == 'P' || c1 == 'P') n1 = 2;
    if (c0 == 'T' || c1 == 'T') n1 = 3;
    if (c1 == 'u') n1 = 4;
    if (c0 == 'E' || c1 == 'E') n1 = 5;
    if (c0 == 'H' || c1 == 'H') n1 = 6;
    if (c0 == 'L' || c1 == 'L') n1 = 7;
    if (c0 == 'X' || c1 == 'X') n1 = 8;
    if (c0 == 'Y' || c1 == 'Y') n1 = 9;
    if (c0 == 'M' || c1 == 'M') n1 = 10;
    if (n1 == -1) return fail("init: unknown boundary parameter " + s);
    long m10, m11, m20, m21;
    double bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
    if (yn == 0) {
      std::string a;
      if (!tok(a)) return false;
      m10 = a[0] == 'x' ? X - 1 + atol(a.c_str() + 1) : atol(a.c_str());
      if (!tok(a)) return false;
      m11 = a[0] == 'x' ? X - 1 + atol(a.c_str() + 1) : atol(a.c_str());
      if (!tok(a)) return false;
      m20 = a[0] == 'y' ? Y - 1 + atol(a.c_str() + 1) : atol(a.c_str());
      if (!tok(a)) return false;
      m21 = a[0] == 'y' ? Y - 1 + atol(a.c_str() + 1) : atol(a.c_str());
    } else {
      for (int n2 = 0; n2 < 4; n2++) {
        bx[n2] = frac(xsize);
        by[n2] = frac(ysize);
      }
      m10 = 0;
      m11 = X - 1;
      m20 = 0;
      m21 = Y - 1;
    }
    const double cnst = dtok();
    // up to three neighbour terms: coefficient, then the shift (dm1, dm2);
    // an 'x'/'y' prefix on dm1 couples Vx to Vy (-1) or Vy to Vx (+1); a
    // 'c' coefficient marks a continuity-type BC (1.1e30, movemart.c:863)
    double kf[3];
    long nsh[3] = {0, 0, 0};
    int ncomp[3] = {0, 0, 0};
    for (int q = 0; q < 3; q++) {
      std::string a;
      if (!tok(a)) return false;
      kf[q] = (q == 0 && a[0] == 'c') ? 1.1e+30 : atof(a.c_str());
      if (kf[q] != 0) {
        if (!tok(a)) return false;
        nsh[q] = atol(a.c_str()) * Y;
        if (a[0] == 'x') {
          ncomp[q] = -1;
          nsh[q] = atol(a.c_str() + 1) * Y;
        }
        if (a[0] == 'y') {
          ncomp[q] = 1;
          nsh[q] = atol(a.c_str() + 1) * Y;
        }
        if (!tok(a)) return false;
        nsh[q] += atol(a.c_str());
      }
    }
    if (!err.empty()) return false;
    if (m10 < 0 || m20 < 0 || m11 > X - 1 || m21 > Y - 1) return fail("init: boundary box outside the grid");
    // grid lines one past the last: the reference's arrays are longer than
    // the grid and zero there (BSS), which is what its staggered
    // coordinates read on the last line
    auto gxe = [&](long m) { return m < X ? gx[m] : 0.0; };
    auto gye = [&](long m) { return m < Y ? gy[m] : 0.0; };
    for (long m1 = m10; m1 <= m11; m1++)
      for (long m2 = m20; m2 <= m21; m2++) {
        const long m4 = m1 * Y + m2;
        double x = gx[m1] / xsize, y = gy[m2] / ysize;
        long m5 = 0;
        // inside test of an irregular rectangle (edges 0-1 left, 2-3 right)
        auto in_rect = [&](double xx, double yy) {
          const double x1 = bx[0] + (yy - by[0]) * (bx[1] - bx[0]) / (by[1] - by[0]);
          const double x2 = bx[2] + (yy - by[2]) * (bx[3] - bx[2]) / (by[3] - by[2]);
          if (!(xx >= x1 && xx <= x2)) return false;
          const double y1 = by[0] + (xx - bx[0]) * (by[2] - by[0]) / (bx[2] - bx[0]);
          const double y2 = by[1] + (xx - bx[1]) * (by[3] - by[1]) / (bx[3] - bx[1]);
          return yy >= y1 && yy <= y2;
        };
        // circular sector: relative radius and angle, or false
        auto in_sector = [&](double xx, double yy, double &rel, double &ang) {
          const double x1 = xx - bx[0], y1 = (yy - by[0]) * ysize / xsize;
          double x4 = bx[2] - pow(x1 * x1 + y1 * y1, 0.5);
          if (!(x4 >= 0)) return false;
          const double x2 = xx - bx[1], y2 = (yy - by[1]) * ysize / xsize;
          const double y4 = pow(x2 * x2 + y2 * y2, 0.5) - by[2] * ysize / xsize;
          if (!(y4 >= 0)) return false;
          double x3 = x1, y3 = y1;
          if (by[3] < 0) {
            x3 = x2;
            y3 = y2;
          }
          double ival = pow(x3 * x3 + y3 * y3, 0.5);
          if (!ival) ival = 1.0;
          ival = x3 / ival;
          ival = asin(fabs(ival)) / pival * 180.0;
          if (x3 >= 0 && y3 > 0) ival = 180.0 - ival;
          if (x3 < 0 && y3 > 0) ival += 180.0;
          if (x3 < 0 && y3 <= 0) ival = 360.0 - ival;
          if (!(ival >= fabs(bx[3]) && ival <= fabs(by[3]))) return false;
          x4 += y4;
          if (!x4) x4 = 1.0;
          rel = y4 / x4;
          ang = ival;
          return true;
        };
        auto set_row = [&](long unknown, double c, bool comp_shift, int stride) {
          const long b = add_row(c);
          bondm[unknown] = b;
          for (int q = 0; q < 3; q++)
            if (kf[q]) {
              bondn[3 * b + q] = unknown + nsh[q] * stride + (comp_shift ? ncomp[q] : 0) + 1;
              bondv[4 * b + q + 1] = kf[q];
            }
        };
        switch (n1) {
          case 0:
          case 1: {
            if (n1 == 0)
              y = (gy[m2] + gye(m2 + 1)) / 2.0 / ysize;
            else
              x = (gx[m1] + gxe(m1 + 1)) / 2.0 / xsize;
            m5 = 3 * m4 + 1 + n1;
            std::vector<double> &v = n1 == 0 ? vx : vy;
            if (yn == 0 || (yn == 1 && in_rect(x, y))) {
              v[m4] = cnst;
              set_row(m5, cnst, true, 3);
            }
            double rel, ang;
            if (yn == 2 && in_sector(x, y, rel, ang)) {
              const double t = n1 == 0 ? cos(ang / 180.0 * pival) : sin(ang / 180.0 * pival);
              const double c = t * (cnst * rel + kf[0] * (1.0 - rel));
              const long b = add_row(c);
              bondm[m5] = b;
              v[m4] = c;
            }
            break;
          }
          case 2:
            x = (gx[m1] + gxe(m1 + 1)) / 2.0 / xsize;
            y = (gy[m2] + gye(m2 + 1)) / 2.0 / ysize;
            m5 = 3 * m4;
            if (yn == 0 || (yn == 1 && in_rect(x, y))) set_row(m5, cnst, false, 3);
            break;
          case 3:
            m5 = N3 + m4;
            if (yn == 0 || (yn == 1 && in_rect(x, y))) {
              const bool from_t = cnst <= 0 && !kf[0] && !kf[1] && !kf[2];
              set_row(m5, from_t ? tk[m4] : cnst, false, 1);
            }
            break;
          case 4:  // viscosity of the node (mu[], not modelled)
          case 5:  // erosion surface (ep[], not modelled)
          case 6:  // hydration surfaces (ep[], not modelled)
          case 7:
            break;
          case 8: {  // grid lines in x (in2mart.c:900-910)
            double iv = (double)(m11 - m10);
            if (!iv) iv = 1.0;
            const double iv1 = (double)(m1 - m10);
            if (m1 < 1) return fail("init: X grid-line definition from line 0");
            gx[m1] = gx[m1 - 1] + cnst + (kf[0] - cnst) * iv1 / iv;
            if (kf[1]) gx[m1] = gx[m1 - 1] + exp(log(cnst) + log(kf[1] / cnst) * iv1 / iv);
            break;
          }
          case 9: {
            double iv = (double)(m21 - m20);
            if (!iv) iv = 1.0;
            const double iv1 = (double)(m2 - m20);
            if (m2 < 1) return fail("init: Y grid-line definition from line 0");
            gy[m2] = gy[m2 - 1] + cnst + (kf[0] - cnst) * iv1 / iv;
            if (kf[1]) gy[m2] = gy[m2 - 1] + exp(log(cnst) + log(kf[1] / cnst) * iv1 / iv);
            break;
          }
          case 10: {  // a marker set in the cell (in2mart.c:929-974)
            if (m1 > X - 2 || m2 > Y - 2 || nsh[0] < 1 || nsh[1] < 1) return fail("init: bad cell marker set");
            const double dx = (gx[m1 + 1] - gx[m1]) / (double)(nsh[0]), dy = (gy[m2 + 1] - gy[m2]) / (double)(nsh[1]);
            double xx = gx[m1] + dx / 2.0;
            for (long a = 0; a < nsh[0]; a++) {
              double yy = gy[m2] + dy / 2.0;
              for (long b = 0; b < nsh[1]; b++) {
                float fx = (float)(xx), fy = (float)(yy);
                if (cnst > 0 && kf[0] > 0)
                  fx += (float)(rnd.next() % ((int)(cnst)*2 + 1) - (int)(cnst)) / ((float)(cnst)) * (float)(dx * kf[0]);
                if (cnst > 0 && kf[1] > 0)
                  fy += (float)(rnd.next() % ((int)(cnst)*2 + 1) - (int)(cnst)) / ((float)(cnst)) * (float)(dy * kf[1]);
                mx.push_back(fx);
                my.push_back(fy);
                mt.push_back(0);
                yy += dy;
              }
              xx += dx;
            }
            break;
          }
        }
      }
    if (!tok(s)) return false;
  }
  // first and last grid lines reset (in2mart.c:981-984)
  gx[0] = 0;
  gx[X - 1] = xsize;
  gy[0] = 0;
  gy[Y - 1] = ysize;
  return true;
}

// in2mart.c:985-1196: rock bodies on the markers up to '~'
bool Gen::bodies() {
  const double pival = 2.000 * asin(1.000);
  std::string s;
  if (!tok(s)) return false;
  while (s[0] != '~') {
    const int yn = atoi(s.c_str());
    std::string a;
    if (!tok(a)) return false;
    int n5 = 0, n1 = atoi(a.c_str());
    const char a0 = a[0], a1 = a.size() > 1 ? a[1] : 0;
    if (a0 == 'i' && a1 != 'a') {
      n1 = atoi(a.c_str() + 1);
      n5 = 1;
    }
    if (a0 == 'a' && a1 != 'i') {
      n1 = atoi(a.c_str() + 1);
      n5 = 2;
    }
    if ((a0 == 'i' && a1 == 'a') || (a0 == 'a' && a1 == 'i')) {
      n1 = atoi(a.c_str() + 2);
      n5 = 3;
    }
    if (n5 & 1) return fail("init: immobile rock bodies are not supported");
    double bx[4], by[4];
    for (int n2 = 0; n2 < 4; n2++) {
      bx[n2] = frac(xsize);
      by[n2] = frac(ysize);
    }
    if (!tok(s)) return false;  // the next entry's first token
    if (!err.empty()) return false;
    if (yn == 1) {  // markers added in the rectangle (in2mart.c:1021-1060)
      const double dx = xstpx / (double)(mnumx), dy = ystpy / (double)(mnumy);
      for (long m1 = (long)(bx[0] * (double)(X - 1)); m1 < (long)(bx[3] * (double)(X - 1) + 1.0); m1++)
        for (long m2 = (long)(by[0] * (double)(Y - 1)); m2 < (long)(by[3] * (double)(Y - 1) + 1.0); m2++) {
          double x = (double)(m1)*xstpx + dx / 2.0;
          for (long m4 = 0; m4 < mnumx; m4++) {
            double y = (double)(m2)*ystpy + dy / 2.0;
            for (long m5 = 0; m5 < mnumy; m5++) {
              float fx = (float)(x), fy = (float)(y);
              const signed char t = (signed char)n1;
              if (nonstab > 0 && t < 100) {
                fx += (float)(rnd.next() % (nonstab * 2 + 1) - nonstab) / ((float)(nonstab)) * (float)(xnonstab);
                fy += (float)(rnd.next() % (nonstab * 2 + 1) - nonstab) / ((float)(nonstab)) * (float)(ynonstab);
              }
              mx.push_back(fx);
              my.push_back(fy);
              mt.push_back(t);
              y += dy;
            }
            x += dx;
          }
        }
      continue;
    }
    // set the body on the existing markers; 'a' appends a copy instead
    const size_t nm = mx.size();
    for (size_t i = 0; i < nm; i++) {
      bool inside = false;
      if (yn < 3) {
        double y = (double)(my[i]) / ysize - by[0];
        double x = xsize * (bx[0] + y * (bx[1] - bx[0]) / (by[1] - by[0]));
        if ((double)(mx[i]) >= x) {
          y = (double)(my[i]) / ysize - by[2];
          x = xsize * (bx[2] + y * (bx[3] - bx[2]) / (by[3] - by[2]));
          if ((double)(mx[i]) <= x) {
            x = (double)(mx[i]) / xsize - bx[0];
            y = ysize * (by[0] + x * (by[2] - by[0]) / (bx[2] - bx[0]));
            if ((double)(my[i]) >= y) {
              x = (double)(mx[i]) / xsize - bx[1];
              y = ysize * (by[1] + x * (by[3] - by[1]) / (bx[3] - bx[1]));
              inside = (double)(my[i]) <= y;
            }
          }
        }
      } else {
        const double x = (double)(mx[i]) / xsize, y = (double)(my[i]) / ysize;
        const double x1 = x - bx[0], y1 = (y - by[0]) * ysize / xsize;
        if ((x1 * x1 + y1 * y1) <= (bx[2] * bx[2])) {
          const double x2 = x - bx[1], y2 = (y - by[1]) * ysize / xsize;
          if ((x2 * x2 + y2 * y2) >= (by[2] * by[2] * ysize * ysize / xsize / xsize)) {
            double xx = x1, yy = y1;
            if (by[3] < 0) {
              xx = x2;
              yy = y2;
            }
            double ival = pow(xx * xx + yy * yy, 0.5);
            if (!ival) ival = 1.0;
            ival = xx / ival;
            ival = asin(fabs(ival)) / pival * 180.0;
            if (xx >= 0 && yy > 0) ival = 180.0 - ival;
            if (xx < 0 && yy > 0) ival += 180.0;
            if (xx < 0 && yy <= 0) ival = 360.0 - ival;
            inside = ival >= fabs(bx[3]) && ival <= fabs(by[3]);
          }
        }
      }
      if (!inside) continue;
      const bool set = yn == 0 || yn == 3, add = yn == 2 || yn == 4;
      if (n5 == 0) {
        if (set) mt[i] = (signed char)n1;
        if (add) mt[i] = (signed char)(mt[i] + n1);
      } else {  // n5 == 2: an additional marker at the same place
        mx.push_back(mx[i]);
        my.push_back(my[i]);
        mt.push_back(set ? (signed char)n1 : (signed char)(mt[i] + n1));
      }
    }
  }
  // sinusoidal perturbation (in2mart.c:1197-1217)
  if (nonstab < 0)
    for (size_t i = 0; i < mx.size(); i++)
      if (mt[i] < 100) {
        if (nonstab == -1) {
          my[i] += (float)(wamp * sin(2.0 * pival / wlen * ((double)(mx[i]) - xsize * 0.5 + wlen * 0.25)));
        } else {
          double ival = 0;
          if ((double)(my[i]) > wymin && (double)(my[i]) <= wymid) ival = wamp * ((double)(my[i]) - wymin) / (wymid - wymin);
          if ((double)(my[i]) > wymid && (double)(my[i]) < wymax) ival = wamp * ((double)(my[i]) - wymid) / (wymax - wymid);
          my[i] += (float)(ival * sin(2.0 * pival / wlen * ((double)(mx[i]) - xsize * 0.5 + wlen * 0.25)));
        }
      }
  return true;
}

// in2mart.c:1245-1391: the initial temperature field up to '~'
bool Gen::temperature() {
  const double pival = 2.000 * asin(1.000);
  std::string s;
  if (!tok(s)) return false;
  while (s[0] != '~') {
    const int yn = atoi(s.c_str());
    double bx[4], by[4], t[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int n2 = 0; n2 < 4; n2++) {
      bx[n2] = frac(xsize);
      by[n2] = frac(ysize);
    }
    for (int k = 0; k < 4; k++) t[k] = dtok();
    if (yn == 4 || yn == 5 || yn == 6)
      for (int k = 4; k < 7; k++) t[k] = dtok();
    if (!tok(s)) return false;
    if (!err.empty()) return false;
    if (yn < 0 || yn > 7) return fail("init: unknown temperature distribution");
    const double dh0 = rock[0].dh;
    for (long m1 = 0; m1 < X; m1++)
      for (long m2 = 0; m2 < Y; m2++) {
        const long m3 = m1 * Y + m2;
        double x = gx[m1] / xsize, y = gy[m2] / ysize;
        if (yn != 3 && yn != 7) {
          const double x1 = bx[0] + (y - by[0]) * (bx[1] - bx[0]) / (by[1] - by[0]);
          const double x2 = bx[2] + (y - by[2]) * (bx[3] - bx[2]) / (by[3] - by[2]);
          if (!(x >= x1 && x <= x2)) continue;
          const double y1 = by[0] + (x - bx[0]) * (by[2] - by[0]) / (bx[2] - bx[0]);
          const double y2 = by[1] + (x - bx[1]) * (by[3] - by[1]) / (bx[3] - bx[1]);
          if (!(y >= y1 && y <= y2)) continue;
          x = (x - x1) / (x2 - x1);
          y = (y - y1) / (y2 - y1);
          if (yn == 0) tk[m3] = (t[0] + y * (t[1] - t[0])) * (1.0 - x) + (t[2] + y * (t[3] - t[2])) * x;
          if (yn == 1) {
            const double ival = y * (t[1] - t[0]);
            const double akf = exp(-dh0 / 8.314 / t[0]), bkf = exp(-dh0 / 8.314 / t[2]) - akf;
            tk[m3] = ival - dh0 / 8.314 / log(akf + bkf * x);
          }
          if (yn == 2) {
            const double ival = x * (t[2] - t[0]);
            const double akf = exp(-dh0 / 8.314 / t[0]), bkf = exp(-dh0 / 8.314 / t[1]) - akf;
            tk[m3] = ival - dh0 / 8.314 / log(akf + bkf * y);
          }
          if (yn == 4 || yn == 5 || yn == 6) {
            const double ival = (gy[m2] - y1 * ysize) / 2.0 / pow(t[6] * (t[4] + (t[5] - t[4]) * x) * 3.15576e+7, 0.5);
            const double tc = (t[1] + (t[3] - t[1]) * x) - (1.0 - erf(ival)) * ((t[1] - t[0]) * (1.0 - x) + (t[3] - t[2]) * x);
            if (yn == 4) tk[m3] = tc;
            if (yn == 5 || yn == 6) {
              const double tb = (t[0] + y * (t[1] - t[0])) * (1.0 - x) + (t[2] + y * (t[3] - t[2])) * x;
              tk[m3] = yn == 5 ? tc * (1.0 - x) + tb * x : tc * x + tb * (1.0 - x);
            }
          }
        } else if (yn == 3) {
          const double x1 = x - bx[0], y1 = (y - by[0]) * ysize / xsize;
          double x4 = bx[2] - pow(x1 * x1 + y1 * y1, 0.5);
          if (!(x4 >= 0)) continue;
          const double x2 = x - bx[1], y2 = (y - by[1]) * ysize / xsize;
          const double y4 = pow(x2 * x2 + y2 * y2, 0.5) - by[2] * ysize / xsize;
          if (!(y4 >= 0)) continue;
          double x3 = x1, y3 = y1;
          if (by[3] < 0) {
            x3 = x2;
            y3 = y2;
          }
          double ival = pow(x3 * x3 + y3 * y3, 0.5);
          if (!ival) ival = 1.0;
          ival = x3 / ival;
          ival = asin(fabs(ival)) / pival * 180.0;
          if (x3 >= 0 && y3 > 0) ival = 180.0 - ival;
          if (x3 < 0 && y3 > 0) ival += 180.0;
          if (x3 < 0 && y3 <= 0) ival = 360.0 - ival;
          if (!(ival >= fabs(bx[3]) && ival <= fabs(by[3]))) continue;
          x4 += y4;
          if (!x4) x4 = 1.0;
          x = y4 / x4;
          x4 = fabs(by[3]) - fabs(bx[3]);
          if (!x4) x4 = 1.0;
          y = (ival - fabs(bx[3])) / x4;
          tk[m3] = (t[0] * y + t[1] * (1.0 - y)) * x + (t[2] * y + t[3] * (1.0 - y)) * (1.0 - x);
        } else {  // yn == 7: decaying impact anomaly (Monteux et al. 2007)
          const double x1 = x - bx[0], y1 = (y - by[0]) * ysize / xsize;
          const double x4 = pow(x1 * x1 + y1 * y1, 0.5);
          if (!(x4 > bx[2])) continue;
          const double x2 = x - bx[1], y2 = (y - by[1]) * ysize / xsize;
          const double y4 = pow(x2 * x2 + y2 * y2, 0.5);
          if (y4 > bx[2]) tk[m3] += t[0] * pow((bx[2] / y4), 4.400);
        }
      }
  }
  return true;
}

// in2mart.c:1393-1453: porosity, marker temperatures (allintert,
// markmart.c:1687-1745), marktmax
void Gen::finish() {
  const size_t M = mx.size();
  mpor.assign(M, 0.0f);
  mk.assign(M, 0.0f);
  mtmax.assign(M, 0.0f);
  for (size_t i = 0; i < M; i++) {
    if (mt[i] == 0 || mt[i] >= 7) mpor[i] = 0.000f;
    if (mt[i] == 5 || mt[i] == 6) mpor[i] = (float)(por_init);
  }
  for (size_t i = 0; i < M; i++) {
    if (!(mx[i] > 0 && my[i] > 0 && (double)(mx[i]) < xsize && (double)(my[i]) < ysize)) continue;
    double x = (double)mx[i], y = (double)my[i];
    if (x < 0) x = 0; else if (x > xsize) x = xsize;
    if (y < 0) y = 0; else if (y > ysize) y = ysize;
    // m1serch/m2serch: the largest m <= n-2 with g[m] <= x (bisection)
    auto srch = [](const std::vector<double> &g, double v) {
      long m0 = 0, m1 = (long)g.size() - 1, m;
      do {
        m = (m0 + m1) / 2;
        if (g[m] > v)
          m1 = m;
        else
          m0 = m;
      } while ((m1 - m0) > 1);
      if (m0 > (long)g.size() - 2) m0 = (long)g.size() - 2;
      return m0;
    };
    const long m10 = srch(gx, x), m20 = srch(gy, y);
    // fdweight n = 1 interpolation weights, m1 outer and m2 inner, from zero
    const double cx5 = gx[m10] - x, cx4 = gx[m10 + 1] - x, cx3 = gx[m10 + 1] - gx[m10];
    const double wx1 = -1.0 * cx5 * 1.0 / cx3, wx0 = cx4 * 1.0 / cx3;
    const double cy5 = gy[m20] - y, cy4 = gy[m20 + 1] - y, cy3 = gy[m20 + 1] - gy[m20];
    const double wy1 = -1.0 * cy5 * 1.0 / cy3, wy0 = cy4 * 1.0 / cy3;
    const long m = m10 * Y + m20;
    double e = 0;
    e += (wx0 * wy0) * tk[m];
    e += (wx0 * wy1) * tk[m + 1];
    e += (wx1 * wy0) * tk[m + Y];
    e += (wx1 * wy1) * tk[m + Y + 1];
    mk[i] = (float)(e);
    if (mt[i] != 0) mtmax[i] = mk[i];
  }
}

bool Gen::run(const char *path) {
  if (!tk_in.load(path)) return fail(std::string("init: cannot read ") + path);
  return header() && markers() && rocks() && bcs() && bodies() && temperature() && (finish(), true);
}

}  // namespace

extern "C" {

int i2e_init_read(const char *init_path, i2e_init_info *info) {
  if (!init_path || !info) return I2E_ERR_ARG;
  Gen g;
  if (!g.run(init_path)) {
    i2e::set_error("%s", g.err.c_str());
    return g.err.find("not supported") != std::string::npos ? I2E_ERR_UNSUPPORTED : I2E_ERR_ARG;
  }
  memset(info, 0, sizeof *info);
  info->xnumx = g.X;
  info->ynumy = g.Y;
  info->mnumx = g.mnumx;
  info->mnumy = g.mnumy;
  info->marknum = (int64_t)g.mx.size();
  info->bondnum = (int64_t)(g.bondv.size() / 4);
  info->rocknum = g.rocknum;
  info->nonstab = g.nonstab;
  info->timesum = g.timesum;
  return I2E_OK;
}

int i2e_init_generate(const char *init_path, i2e_params *p, i2e_host_state *s) {
  if (!init_path || !p || !s) return I2E_ERR_ARG;
  Gen g;
  if (!g.run(init_path)) {
    i2e::set_error("%s", g.err.c_str());
    return g.err.find("not supported") != std::string::npos ? I2E_ERR_UNSUPPORTED : I2E_ERR_ARG;
  }
  const int64_t N = g.X * g.Y, M = (int64_t)g.mx.size(), B = (int64_t)(g.bondv.size() / 4);
  if (s->marknum != M || s->bondnum != B || !s->gx || !s->gy || !s->markx || !s->bondm) {
    i2e::set_error("init: the state must hold exactly %lld markers and %lld BC rows (i2e_init_read)", (long long)M,
                   (long long)B);
    return I2E_ERR_ARG;
  }
  p->xnumx = g.X;
  p->ynumy = g.Y;
  p->mnumx = g.mnumx;
  p->mnumy = g.mnumy;
  p->xsize = g.xsize;
  p->ysize = g.ysize;
  p->gykoef = g.gykoef;
  p->tmp_ambient = g.tmp_ambient;
  p->al2627_init = g.al2627;
  p->fe6056_init = g.fe6056;
  p->por_init = g.por_init;
  p->ntypes = g.rocknum;
  for (int t = 0; t < I2E_MAXTYPES; t++) {
    p->rock[t] = g.rock[t];
    p->immobile[t] = g.markim[t];
  }
  memcpy(s->gx, g.gx.data(), sizeof(double) * g.X);
  memcpy(s->gy, g.gy.data(), sizeof(double) * g.Y);
  double *zero[] = {s->pr, s->gp, s->ga, s->exx, s->eyy, s->exy, s->sxx, s->syy, s->sxy,
                    s->nu, s->nd, s->ro, s->cp, s->kt, s->ht, s->et};
  for (double *z : zero)
    if (z) memset(z, 0, sizeof(double) * N);
  memcpy(s->tk, g.tk.data(), sizeof(double) * N);
  memcpy(s->vx, g.vx.data(), sizeof(double) * N);
  memcpy(s->vy, g.vy.data(), sizeof(double) * N);
  memcpy(s->markx, g.mx.data(), sizeof(float) * M);
  memcpy(s->marky, g.my.data(), sizeof(float) * M);
  memcpy(s->markk, g.mk.data(), sizeof(float) * M);
  memcpy(s->markpor, g.mpor.data(), sizeof(float) * M);
  memcpy(s->marktmax, g.mtmax.data(), sizeof(float) * M);
  memset(s->markv, 0, sizeof(float) * M);
  memset(s->markp, 0, sizeof(float) * M);
  memcpy(s->markt, g.mt.data(), M);
  memcpy(s->bondm, g.bondm.data(), sizeof(int64_t) * 4 * N);
  memcpy(s->bondn, g.bondn.data(), sizeof(int64_t) * 3 * B);
  memcpy(s->bondv, g.bondv.data(), sizeof(double) * 4 * B);
  s->timestep = 0;
  s->timesum = g.timesum;
  s->gxkoef = g.gxkoef;
  return I2E_OK;
}

int i2e_mode_read(const char *mode_path, int file_index, i2e_params *p, int *cyc0max) {
  if (!mode_path || !p || file_index < 0) return I2E_ERR_ARG;
  Tokens t;
  if (!t.load(mode_path)) {
    i2e::set_error("mode: cannot read %s", mode_path);
    return I2E_ERR_ARG;
  }
  std::string s;
  bool ok = true;
  auto tok = [&]() {
    if (!t.next(s)) ok = false;
    return ok ? s : std::string("0");
  };
  // loadconf (loadmart.c:2-105): the input file and its type, then one line
  // per output file (name, type, cyc0max, maxxystep, maxtkstep, maxtmstep in
  // years) up to '~', then the GENERAL, STOKES and HEAT parameters
  tok();
  tok();
  int n = 0, found = 0;
  for (std::string name = tok(); ok && name[0] != '~'; name = tok(), n++) {
    tok();
    const int cyc = atoi(tok().c_str());
    const double xy = atof(tok().c_str()), tkst = atof(tok().c_str()), tm = atof(tok().c_str()) * 3.15576e+7;
    if (n == file_index) {
      found = 1;
      if (cyc0max) *cyc0max = cyc;
      p->maxxystep = xy;
      p->maxtkstep = tkst;
      p->maxtmstep = tm;
    }
  }
  tok();  // printmod
  tok();  // timedir
  tok();  // movemod
  tok();  // tempmod
  const int markmod = atoi(tok().c_str());
  tok();  // ratemod
  tok();  // gridmod
  tok();  // intermod
  tok();  // intermod1
  const int outgrid = atoi(tok().c_str()), densimod = atoi(tok().c_str());
  tok();  // DIVVMIN
  tok();  // STOKSMIN
  tok();  // stoksmod
  tok();  // presmod
  tok();  // stoksfd
  const double nubeg = atof(tok().c_str()), nuend = atof(tok().c_str()), nucontr = atof(tok().c_str());
  tok();  // hidry
  tok();  // hidrl
  const double strmin = atof(tok().c_str());
  tok();  // strmax
  tok();  // viscmod
  tok();  // HEATMIN
  const int heatmod = atoi(tok().c_str());
  tok();  // heatfd
  const double heatdif = atof(tok().c_str());
  const double frictyn = atof(tok().c_str()) * 3.15576e+7, adiabyn = atof(tok().c_str()) * 3.15576e+7;
  if (!ok) {
    i2e::set_error("mode: %s ends early", mode_path);
    return I2E_ERR_ARG;
  }
  if (!found) {
    i2e::set_error("mode: no output file %d (the file lists %d)", file_index, n);
    return I2E_ERR_ARG;
  }
  p->markmod = markmod;
  p->outgrid = outgrid;
  p->densimod = densimod;
  p->nubeg = nubeg;
  p->nuend = nuend;
  p->nucontr = nucontr;
  p->strmin = strmin;
  p->heatmod = heatmod;
  p->heatdif = heatdif;
  p->frictyn = frictyn;
  p->adiabyn = adiabyn;
  return I2E_OK;
}

}  // extern "C"
