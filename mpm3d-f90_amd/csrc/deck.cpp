// deck.cpp -- reader for the MPM3D .mpm input deck.
//
// Reproduces the reference's free-format reader and deck semantics so that
// the problem handed to the device is bit-identical to the state the
// reference builds before its first step:
//   * tokenizer: FFI.f90:177-253 (ReadLine: tabs -> blanks, '!' comments,
//     blank/comma separators, tokens truncated to 20 characters),
//     FFI.f90:255-285 (4-character case-insensitive keyword match),
//     FFI.f90:287-299 (a numeric field starts with '-' or a digit);
//   * every real goes through GetReal, which returns a default (FP32) REAL
//     (FFI.f90:110-137), so each deck real is rounded to float first;
//   * keywords: DataIn.f90:115-361; materials :368-532; particle generators
//     :534-865; loads :884-944; velocities :946-1042; EOS :1095-1156;
//     Initial :1211-1267; SetDT :1269-1297; grid: Grid.f90:222-418.
// Keywords that only steer file output are consumed and ignored.  Features
// outside the explicit updated-Lagrangian step (contact, TLMPM, dynamic
// relaxation, Johnson-Cook family, HE burn, EOS 1-3) are rejected with a
// message naming the keyword.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mpm3d_problem.h"
#include "mpm3d_hip.h"

namespace {

struct DeckError {
  std::string msg;
};

// The line as FFI's ReadLine leaves it in sss (FFI.f90:191-233), which
// GetString hands out (the deck title): tabs to blanks, the comment cut,
// leading blanks removed, every blank followed by a blank or a comma
// removed, then every blank right after a comma removed; 256 characters,
// blank-padded.
static std::string ffi_line(const std::string& line) {
  std::string s = line;
  for (auto& c : s)
    if (c == '\t') c = ' ';
  const size_t bang = s.find('!');
  if (bang != std::string::npos) s.resize(bang);
  const size_t a = s.find_first_not_of(' ');
  s = a == std::string::npos ? std::string() : s.substr(a);
  const size_t z = s.find_last_not_of(' ');
  int len = z == std::string::npos ? 0 : (int)z + 1;
  std::string b = s;
  b.resize(258, ' ');
  auto ch = [&](int i) { return b[i - 1]; };   // 1-based, as sss(i:i)
  auto del = [&](int i) {                     // sss(i:len-1) = sss(i+1:len); sss(len:len) = ' '
    b.erase((size_t)i - 1, 1);
    b.push_back(' ');
  };
  for (int i = 1; i < len; i++)
    while (ch(i) == ' ' && (ch(i + 1) == ' ' || ch(i + 1) == ',')) {
      del(i);
      len--;
    }
  for (int i = 1; i < len; i++)
    if (ch(i) == ',' && ch(i + 1) == ' ') {
      del(i + 1);
      len--;
    }
  b.resize(256);
  return b;
}

struct Tokenizer {
  FILE* f = nullptr;
  int line_nb = 0;
  std::vector<std::string> words;  // tokens of the current line
  size_t next = 0;                 // next unread token
  std::string raw;                 // the current line as read

  size_t remaining() const { return words.size() - next; }

  [[noreturn]] void fail(const std::string& m) const {
    throw DeckError{m + " (deck line " + std::to_string(line_nb) + ")"};
  }

  // FFI.f90:177-253.  Returns false at end of file.
  bool read_line() {
    char buf[4096];
    if (!fgets(buf, sizeof buf, f)) return false;
    line_nb++;
    std::string s(buf);
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
    if (s.size() > 256) s.resize(256);  // character(256) sss
    raw = s;
    for (auto& c : s)
      if (c == '\t') c = ' ';
    size_t bang = s.find('!');
    if (bang != std::string::npos) s.resize(bang);
    words.clear();
    next = 0;
    std::string cur;
    for (char c : s) {
      if (c == ' ' || c == ',') {
        if (!cur.empty()) words.push_back(cur);
        cur.clear();
      } else {
        cur.push_back(c);
      }
    }
    if (!cur.empty()) words.push_back(cur);
    for (auto& w : words)
      if (w.size() > 20) w.resize(20);  // character(20) cmd_line(15)
    return true;
  }

  void ensure_word() {
    while (remaining() == 0) {
      if (!read_line()) fail("end of file reached before 'endi'");
    }
  }

  const std::string& take() {
    ensure_word();
    return words[next++];
  }

  static bool is_number(const std::string& w) {
    return !w.empty() && (w[0] == '-' || (w[0] >= '0' && w[0] <= '9'));
  }

  // GetInt (FFI.f90:81-108)
  int get_int() {
    const std::string& w = take();
    if (!is_number(w)) fail("non-numeric field '" + w + "' where an integer is expected");
    char* end = nullptr;
    long v = strtol(w.c_str(), &end, 10);
    if (end == w.c_str()) fail("bad integer '" + w + "'");
    return (int)v;
  }

  // GetReal returns a default REAL: the decimal value rounded once to FP32.
  double get_real() {
    const std::string& w0 = take();
    if (!is_number(w0)) fail("non-numeric field '" + w0 + "' where a real is expected");
    std::string w = w0;
    for (auto& c : w)
      if (c == 'd' || c == 'D') c = 'e';
    char* end = nullptr;
    float v = strtof(w.c_str(), &end);
    if (end == w.c_str()) fail("bad real '" + w0 + "'");
    return (double)v;
  }

  // pcomp over 4 blank-padded characters, case-insensitive (FFI.f90:255-285)
  static bool kw_match(const std::string& w, const char* kw) {
    const int inc = 'A' - 'a';
    const size_t kl = strlen(kw);
    for (size_t i = 0; i < 4; i++) {
      int a = i < w.size() ? (unsigned char)w[i] : ' ';
      int b = i < kl ? (unsigned char)kw[i] : ' ';
      if (a != b && a + inc != b && a != b + inc) return false;
    }
    return true;
  }

  // KeyWord (FFI.f90:140-175); returns 1-based index or -1.
  int keyword(const char* const* kws, int n) {
    ensure_word();
    const std::string& w = words[next];
    for (int i = 0; i < n; i++) {
      if (kw_match(w, kws[i])) {
        next++;
        return i + 1;
      }
    }
    return -1;
  }
};

struct Deck {
  mpm_outputs outs{};   // output options (DataOut.f90:39-59 defaults set in the constructor)
  Deck() {
    memset(outs.title, ' ', 256);
    outs.n_curves = 2;                // engk and engi of particle 1
    outs.curve_option[0] = 6;
    outs.curve_option[1] = 7;
    outs.curve_point[0] = outs.curve_point[1] = 1;
    outs.n_anim = 1;                  // the material number
    outs.anim_option[0] = 3;
  }
  Tokenizer tk;
  int nb_particle = 0, nb_body = 0, nb_mat = 0, nb_component = 1;
  std::vector<mpm_material> mats;
  std::vector<mpm_body> bodies;
  std::vector<double> xp[3], vxp[3], fxp[3], mass;
  double span[6] = {0, 0, 0, 0, 0, 0};
  double dcell = 0, cutoff = 0, dtscale = (double)0.9f, endtime = 0;
  double qlt = 0, damp = 0, bq1 = 1.5, bq2 = (double)0.06f;
  bool dr_damping = false;
  bool smooth_stress = false;   // smoo: SmoothStressbyGrid every step (Spre, read by the ParaView writer)
  bool tlmpm = false;           // TLmp: total-Lagrangian step sequence
  int contact = 0;              // cont: Contact_type (1 lagr: Lagr_NodContact; 2 others: no contact force)
  double fricfa = 0;            // friction coefficient
  int normbody = 0;             // contact normal from both bodies (0), component 1 (1) or 2 (2)
  int n_deto = 0;
  double deto[10][3] = {};
  int fix[6] = {0, 0, 0, 0, 0, 0};
  bool musl = false, usl = false, usf = false, update_method = false;
  bool gimp = false, bspline = false, sgmp = false, quasi_load = false, quasi_damp = false;
  bool gravity = false, uvfp = false;
  int body_counter = 0, par_counter = 0;

  [[noreturn]] void fail(const std::string& m) { tk.fail(m); }

  bool on_off() {
    static const char* kw[] = {"on", "off"};
    int k = tk.keyword(kw, 2);
    if (k == 1) return true;
    if (k == 2) return false;
    fail("switch should be ON/OFF");
  }

  int res_option() {
    static const char* kw[] = {"seqv", "epef", "mat", "pres", "volu", "engk", "engi", "velx", "vely",
                               "velz", "cels", "fail", "sspd", "damg", "stsx", "disx", "disy", "disz"};
    int k = tk.keyword(kw, 18);
    if (k <= 0) fail("result option error");
    return k;
  }

  void need_particles() {
    if (nb_particle == 0) fail("nbmp must be defined in advance");
  }

  mpm_body& new_body() {
    body_counter++;
    if (body_counter > nb_body) fail("more particle sets than 'nbbo' bodies");
    return bodies[body_counter - 1];
  }

  void add_particle(double m, double x, double y, double z) {
    par_counter++;
    if (par_counter > nb_particle) fail("too many particles for 'nbmp'");
    int p = par_counter - 1;
    mass[p] = m;
    xp[0][p] = x;
    xp[1][p] = y;
    xp[2][p] = z;
  }

  int check_mat(int m) {
    if (m > nb_mat || m <= 0) fail("improper material ID");
    return m;
  }

  // SetMaterial (DataIn.f90:368-532)
  void set_material() {
    static const char* kw[] = {"elas", "pla1", "pla2", "john", "sjc", "sjcf", "jcf", "hiex", "null", "dpm", "neoh"};
    if (nb_mat == 0 || nb_particle == 0) fail("nmat/nbmp must be defined before 'material'");
    int i = 0;
    while (i < nb_mat) {
      i = tk.get_int();
      if (i > nb_mat || i <= 0) fail("too many material sets");
      mpm_material& m = mats[i - 1];
      int t = tk.keyword(kw, 11);
      switch (t) {
        case 1:
          m.type = 1;
          m.density = tk.get_real(); m.young = tk.get_real(); m.poisson = tk.get_real();
          break;
        case 2:
          m.type = 2;
          m.density = tk.get_real(); m.young = tk.get_real(); m.poisson = tk.get_real();
          m.yield0 = tk.get_real();
          break;
        case 3:
          m.type = 3;
          m.density = tk.get_real(); m.young = tk.get_real(); m.poisson = tk.get_real();
          m.yield0 = tk.get_real(); m.tangmod = tk.get_real();
          break;
        case 9:
          m.type = 9;
          m.density = tk.get_real(); m.bulk_modulus = tk.get_real(); m.miu = tk.get_real();
          m.wavespd = std::sqrt(m.bulk_modulus / m.density);
          break;
        case 10:
          m.type = 10;
          m.density = tk.get_real(); m.young = tk.get_real(); m.poisson = tk.get_real();
          m.q_fai = tk.get_real(); m.k_fai = tk.get_real(); m.q_psi = tk.get_real(); m.ten_f = tk.get_real();
          break;
        case 11:
          m.type = 11;
          m.density = tk.get_real(); m.young = tk.get_real(); m.poisson = tk.get_real();
          uvfp = true;
          break;
        case 4: case 7:   // john / jcf
          m.type = t;
          m.density = tk.get_real(); m.young = tk.get_real(); m.poisson = tk.get_real();
          m.yield0 = tk.get_real();
          m.b_jc = tk.get_real(); m.n_jc = tk.get_real(); m.c_jc = tk.get_real(); m.m_jc = tk.get_real();
          m.roomt = tk.get_real(); m.melt = tk.get_real(); m.spec_heat = tk.get_real();
          m.epso = tk.get_real();
          if (t == 7) m.epf = tk.get_real();
          break;
        case 5: case 6:   // sjc / sjcf
          m.type = t;
          m.density = tk.get_real(); m.young = tk.get_real(); m.poisson = tk.get_real();
          m.yield0 = tk.get_real();
          m.b_jc = tk.get_real(); m.n_jc = tk.get_real(); m.c_jc = tk.get_real();
          m.epso = tk.get_real();
          if (t == 6) m.epf = tk.get_real();
          break;
        case 8:   // hiex
          m.type = 8;
          m.density = tk.get_real(); m.det_speed = tk.get_real();
          break;
        default:
          fail("invalid material type");
      }
    }
    // CutOff (DataIn.f90:517-524): FP32 literal 1e-5, summed densities / count
    double c = 0.0;
    int count = 0;
    for (int k = 0; k < nb_mat; k++) {
      c = mats[k].density + c;
      count++;
    }
    cutoff = c * (dcell * dcell * dcell) * (double)1e-5f / (double)count;
  }

  // SetParticle (DataIn.f90:534-865)
  void set_particle() {
    static const char* kw[] = {"poin", "bloc", "sphe", "blod", "cicl"};
    need_particles();
    int opt = tk.keyword(kw, 5);
    switch (opt) {
      case 1: {  // point: comID pnum { id mat mass x y z }
        mpm_body& b = new_body();
        b.com = tk.get_int();
        int pnum = tk.get_int();
        b.par_begin = par_counter;
        int mat = 0;
        for (int p = 0; p < pnum; p++) {
          int newi = tk.get_int();
          if (newi < par_counter + 1) fail("particle ID conflict");
          mat = check_mat(tk.get_int());
          double m = tk.get_real();
          double x = tk.get_real(), y = tk.get_real(), z = tk.get_real();
          add_particle(m, x, y, z);
        }
        b.par_end = b.par_begin + pnum;
        b.mat = mat;
        break;
      }
      case 2: case 4: {  // block / block with cylindrical hole
        mpm_body& b = new_body();
        b.com = tk.get_int();
        b.mat = check_mat(tk.get_int());
        b.par_begin = par_counter;
        double pm = tk.get_real();
        double dp = tk.get_real();
        pm = pm * (dp * dp * dp);
        double ox = tk.get_real(), oy = tk.get_real(), oz = tk.get_real();
        ox = ox + dp * 0.5;
        oy = oy + dp * 0.5;
        oz = oz + dp * 0.5;
        int nx = tk.get_int(), ny = tk.get_int(), nz = tk.get_int();
        double oox = 0, ooy = 0, r2 = 0;
        if (opt == 4) {
          oox = tk.get_real(); ooy = tk.get_real(); (void)tk.get_real();
          int nr = tk.get_int();
          double r = (double)nr * dp;
          r2 = r * r;
        }
        for (int iz = 1; iz <= nz; iz++)
          for (int iy = 1; iy <= ny; iy++)
            for (int ix = 1; ix <= nx; ix++) {
              double tx = (double)(ix - 1) * dp + ox;
              double ty = (double)(iy - 1) * dp + oy;
              double tz = (double)(iz - 1) * dp + oz;
              if (opt == 4) {
                double ddx = tx - oox, ddy = ty - ooy;
                if (!(ddx * ddx + ddy * ddy >= r2)) continue;
              }
              add_particle(pm, tx, ty, tz);
            }
        b.par_end = par_counter;
        break;
      }
      case 3: {  // sphere
        mpm_body& b = new_body();
        b.com = tk.get_int();
        b.mat = check_mat(tk.get_int());
        b.par_begin = par_counter;
        double pm = tk.get_real();
        double dp = tk.get_real();
        pm = pm * (dp * dp * dp);
        double ox = tk.get_real(), oy = tk.get_real(), oz = tk.get_real();
        int n = (int)tk.get_real();
        double r = (double)n * dp, r2 = r * r;
        for (int ix = 1 - n; ix <= n; ix++)
          for (int iy = 1 - n; iy <= n; iy++)
            for (int iz = 1 - n; iz <= n; iz++) {
              double tx = (double)(ix - 1) * dp + ox + 0.5 * dp;
              double ty = (double)(iy - 1) * dp + oy + 0.5 * dp;
              double tz = (double)(iz - 1) * dp + oz + 0.5 * dp;
              double a = tx - ox, bb = ty - oy, c = tz - oz;
              if (a * a + bb * bb + c * c < r2) add_particle(pm, tx, ty, tz);
            }
        b.par_end = par_counter;
        break;
      }
      case 5: {  // circular plate
        mpm_body& b = new_body();
        b.com = tk.get_int();
        b.mat = check_mat(tk.get_int());
        b.par_begin = par_counter;
        double pm = tk.get_real();
        double dp = tk.get_real();
        pm = pm * (dp * dp * dp);
        double ox = tk.get_real(), oy = tk.get_real(), oz = tk.get_real();
        int n = tk.get_int();
        double r = (double)n * dp, r2 = r * r;
        for (int ix = 1 - n; ix <= n; ix++)
          for (int iy = 1 - n; iy <= n; iy++) {
            double tx = (double)(ix - 1) * dp + ox + 0.5 * dp;
            double ty = (double)(iy - 1) * dp + oy + 0.5 * dp;
            double tz = oz + dp * 0.5;
            double a = tx - ox, bb = ty - oy;
            if (a * a + bb * bb < r2) add_particle(pm, tx, ty, tz);
          }
        b.par_end = par_counter;
        break;
      }
      default:
        fail("particle option should be point/block/sphere/blod/cicl");
    }
  }

  void need_bodies() {
    if (nb_body == 0) fail("nbbo must be defined before loads/velocities");
  }

  int check_body(int b) {
    if (b <= 0 || b > body_counter) fail("body id out of range");
    return b - 1;
  }

  int check_part(int p) {
    if (p <= 0 || p > nb_particle) fail("particle id out of range");
    return p - 1;
  }

  // SetLoad (DataIn.f90:884-944)
  void set_load() {
    static const char* kw[] = {"endl", "part", "body", "grav", "node"};
    need_bodies();
    for (;;) {
      int k = tk.keyword(kw, 5);
      if (k == 1) break;
      if (k == 2) {
        int p = check_part(tk.get_int());
        fxp[0][p] = tk.get_real(); fxp[1][p] = tk.get_real(); fxp[2][p] = tk.get_real();
      } else if (k == 3) {
        int b = check_body(tk.get_int());
        double f0 = tk.get_real(), f1 = tk.get_real(), f2 = tk.get_real();
        for (int i = bodies[b].par_begin; i < bodies[b].par_end; i++) {
          fxp[0][i] = f0; fxp[1][i] = f1; fxp[2][i] = f2;
        }
      } else if (k == 4) {
        gravity = true;
        int b = check_body(tk.get_int());
        bodies[b].grav[0] = tk.get_real(); bodies[b].grav[1] = tk.get_real(); bodies[b].grav[2] = tk.get_real();
      } else if (k == 5) {
        fail("load keyword 'node' is a legacy form; use 'part'");
      } else {
        fail("error in load input");
      }
    }
  }

  // SetVelocity (DataIn.f90:946-1042)
  void set_velocity() {
    static const char* kw[] = {"endv", "node", "body", "ksin", "nuli", "vsin"};
    need_bodies();
    for (;;) {
      int k = tk.keyword(kw, 6);
      if (k == 1) break;
      if (k == 2) {
        int p = check_part(tk.get_int());
        vxp[0][p] = tk.get_real(); vxp[1][p] = tk.get_real(); vxp[2][p] = tk.get_real();
      } else if (k == 3) {
        int b = check_body(tk.get_int());
        double v0 = tk.get_real(), v1 = tk.get_real(), v2 = tk.get_real();
        for (int i = bodies[b].par_begin; i < bodies[b].par_end; i++) {
          vxp[0][i] = v0; vxp[1][i] = v1; vxp[2][i] = v2;
        }
      } else if (k == 6) {
        // v_x = v0*sin(pi*y0/6) with the FP32 literal 3.1415926535 (DataIn.f90:1022-1033)
        double v0 = tk.get_real();
        const double pi32 = (double)3.1415926535f;
        for (int b = 0; b < body_counter; b++)
          for (int p = bodies[b].par_begin; p < bodies[b].par_end; p++) {
            vxp[0][p] = v0 * std::sin(pi32 * xp[1][p] / 6.0);
            vxp[1][p] = 0.0;
            vxp[2][p] = 0.0;
          }
      } else if (k == 4 || k == 5) {
        fail("hard-wired velocity profiles 'ksin'/'nuli' are not supported");
      } else {
        fail("error in velocity input");
      }
    }
  }

  void set_eos() {
    int mset = tk.get_int();
    if (mset <= 0 || mset > nb_mat) fail("seos: material id out of range");
    int et = tk.get_int();
    mpm_material& m = mats[mset - 1];
    m.eos_type = et;
    switch (et) {   // SetEos (DataIn.f90:1095-1155)
      case 1:       // linear polynomial: C0..C6, E0
        for (int k = 0; k < 7; k++) m.ceos[k] = tk.get_real();
        m.ceos[9] = tk.get_real();
        break;
      case 2: {     // Mie-Gruneisen from c0, lambda (s), gamma0, E0
        const double c0 = tk.get_real(), lambda = tk.get_real(), gamma0 = tk.get_real();
        m.ceos[0] = m.density * (c0 * c0);
        m.ceos[1] = m.ceos[0] * (2 * lambda - 1);
        m.ceos[2] = m.ceos[0] * (lambda - 1) * (3 * lambda - 1);
        m.ceos[3] = m.density * gamma0;
        m.ceos[4] = gamma0;
        m.ceos[9] = tk.get_real();
        break;
      }
      case 3:       // JWL: A, B, R1, R2, omega, E0
        for (int k = 0; k < 5; k++) m.ceos[k] = tk.get_real();
        m.ceos[9] = tk.get_real();
        break;
      case 4:       // weakly compressible: c
        m.ceos[0] = tk.get_real();
        break;
      default:
        fail("error in EOS type (seos)");
    }
  }

  void parse() {
    static const char* kw[] = {"endi", "mpm3", "nbmp", "endt", "grid", "spx", "spy", "spz", "dcel", "dtsc", "outt",
                               "rptt", "fixe", "nmat", "mate", "part", "musl", "jaum", "load", "velo", "outr", "curv",
                               "seos", "pt2d", "curx", "deto", "tecp", "bulk", "gimp", "cont", "usf", "usl", "nbco",
                               "nbbo", "para", "drda", "bimp", "sgmp", "smoo", "TLmp", "UVFp", "QuLo", "Damp"};
    for (;;) {
      int key = tk.keyword(kw, 43);
      switch (key) {
        case 1: return;
        case 2: {   // title: GetString hands out the whole line (DataIn.f90:134)
          const std::string t = ffi_line(tk.raw);
          memcpy(outs.title, t.data(), 256);
          outs.title[256] = 0;
          tk.next = tk.words.size();
          break;
        }
        case 3: {
          nb_particle = tk.get_int();
          if (nb_particle <= 0) fail("nbmp must be positive");
          for (int c = 0; c < 3; c++) {
            xp[c].assign(nb_particle, 0.0);
            vxp[c].assign(nb_particle, 0.0);
            fxp[c].assign(nb_particle, 0.0);
          }
          mass.assign(nb_particle, 0.0);
          break;
        }
        case 4: endtime = tk.get_real(); outs.out_time = endtime; break;   // OutTime = EndTime
        case 5: for (int i = 0; i < 6; i++) span[i] = tk.get_real(); break;
        case 6: span[0] = tk.get_real(); span[1] = tk.get_real(); break;
        case 7: span[2] = tk.get_real(); span[3] = tk.get_real(); break;
        case 8: span[4] = tk.get_real(); span[5] = tk.get_real(); break;
        case 9: dcell = tk.get_real(); break;
        case 10:
          dtscale = tk.get_real();
          if (dtscale > 2.5) fail("DTScale too large");
          break;
        case 11: outs.out_time = tk.get_real(); break;
        case 12: outs.report_time = tk.get_real(); break;
        case 13: for (int i = 0; i < 6; i++) fix[i] = tk.get_int(); break;
        case 14:
          nb_mat = tk.get_int();
          if (nb_mat <= 0) fail("nmat must be positive");
          mats.assign(nb_mat, mpm_material{});
          break;
        case 15: set_material(); break;
        case 16: set_particle(); break;
        case 17: musl = on_off(); if (musl) update_method = true; break;
        case 18: (void)on_off(); break;  // jaum: Jaumann rotation is always applied
        case 19: set_load(); break;
        case 20: set_velocity(); break;
        case 21:
          if (outs.n_anim >= MPM_MAX_ANIM) fail("more than 10 'outr' variables");
          outs.anim_option[outs.n_anim++] = res_option();
          break;
        case 22: {
          if (outs.n_curves >= MPM_MAX_CURVES) fail("more than 15 curves");
          const int k = outs.n_curves++;
          outs.curve_option[k] = res_option();
          outs.curve_point[k] = 1;
          if (tk.remaining() > 0) {
            int t = tk.get_int();
            if (t <= 0 || t > nb_particle) fail("invalid curve particle ID");
            outs.curve_point[k] = t;
          }
          break;
        }
        case 23: set_eos(); break;
        case 24:   // pt2d: particles generated from here on outside the box are not plotted
          outs.plot2d = 1;
          outs.skip_first = par_counter;
          for (int i = 0; i < 6; i++) outs.plot2d_box[i] = (double)(float)tk.get_real();   // real plot2d(6)
          break;
        case 25: {   // SetCurX (DataIn.f90:1158-1192): the last body's first particle within DCell/4
          if (par_counter != nb_particle) fail("curx should be used after particle input");
          if (outs.n_curves >= MPM_MAX_CURVES) fail("more than 15 curves");
          const int k = outs.n_curves++;
          outs.curve_option[k] = res_option();
          outs.curve_point[k] = 1;
          double t[3];
          for (int i = 0; i < 3; i++) t[i] = tk.get_real();
          for (size_t b = 0; b < bodies.size(); b++)
            for (int p = bodies[b].par_begin; p < bodies[b].par_end; p++) {
              const double m = std::max({std::fabs(xp[0][p] - t[0]), std::fabs(xp[1][p] - t[1]),
                                         std::fabs(xp[2][p] - t[2])});
              if (m <= 0.25 * dcell) {
                outs.curve_point[k] = p + 1;
                break;
              }
            }
          break;
        }
        case 26:   // deto: a detonation point (DataIn.f90:265-269)
          if (n_deto >= 10) fail("more than 10 detonation points");
          for (int i = 0; i < 3; i++) deto[n_deto][i] = tk.get_real();
          n_deto++;
          break;
        case 27: outs.tecplot = 1; break;
        case 28: bq1 = tk.get_real(); bq2 = tk.get_real(); break;
        case 29: gimp = true; break;
        case 30: {   // Setcontact (DataIn.f90:1348-1393): lagr (Bardenhagen) or others, friction, normal
          static const char* ck[] = {"lagr", "othe"};
          contact = tk.keyword(ck, 2);
          if (contact <= 0) fail("Error set contact parameter (lagr/pena)");
          fricfa = tk.get_real();
          normbody = tk.get_int();
          break;
        }
        case 31: usf = on_off(); if (usf) update_method = true; break;
        case 32: usl = on_off(); if (usl) update_method = true; break;
        case 33:
          // without contact every body maps to component 1 (update_step.f90:56);
          // the other components' grids stay empty
          nb_component = tk.get_int();
          if (nb_component < 1) fail("nbco must be positive");
          break;
        case 34:
          nb_body = tk.get_int();
          if (nb_body <= 0) fail("nbbo must be positive");
          bodies.assign(nb_body, mpm_body{});
          break;
        case 35: outs.paraview = 1; break;
        case 36: dr_damping = true; break;
        case 37: bspline = true; break;
        case 38: sgmp = true; break;
        case 39: smooth_stress = true; break;
        case 40: tlmpm = true; break;
        case 41: uvfp = true; break;
        case 42: quasi_load = true; qlt = tk.get_real(); break;
        case 43: quasi_damp = true; damp = tk.get_real(); break;
        default:
          fail("invalid keyword '" + (tk.remaining() ? tk.words[tk.next] : std::string("?")) + "'");
      }
    }
  }
};

}  // namespace

extern "C" int mpm_deck_read(const char* path, mpm_problem** out, char* errbuf, int errlen) {
  if (!out) return MPM_ERR_ARG;
  *out = nullptr;
  Deck d;
  d.tk.f = fopen(path, "r");
  if (!d.tk.f) {
    if (errbuf && errlen > 0) snprintf(errbuf, errlen, "cannot open deck '%s'", path);
    return MPM_ERR_IO;
  }
  try {
    d.parse();
    fclose(d.tk.f);
    d.tk.f = nullptr;
    if (!d.update_method) d.musl = true;  // DataIn.f90:364
    if (d.par_counter != d.nb_particle) d.fail("fewer particles generated than 'nbmp'");
    if (d.body_counter != d.nb_body) d.fail("fewer particle sets than 'nbbo'");
    if (d.dcell == 0.0) d.fail("DCell must be defined");
    for (int m = 0; m < d.nb_mat; m++)
      if (d.mats[m].type == 0) d.fail("material " + std::to_string(m + 1) + " not defined");
    for (int m = 0; m < d.nb_mat; m++) {
      const mpm_material& mt = d.mats[m];
      const std::string id = "material " + std::to_string(m + 1);
      // seleos is called by sjc, sjcf, jcf, hiex and null (Constitution.f90:239-330)
      const bool needs_eos = mt.type == MPM_MAT_SJC || mt.type == MPM_MAT_SJCF || mt.type == MPM_MAT_JCF ||
                             mt.type == MPM_MAT_HIEX || mt.type == MPM_MAT_NULL;
      if (needs_eos && (mt.eos_type < 1 || mt.eos_type > 4)) d.fail(id + " needs an equation of state ('seos')");
      // eos3 scales by the burn fraction only HE burn sets (for others the
      // reference reads the last HE particle's value)
      if (mt.eos_type == 3 && mt.type != MPM_MAT_HIEX) d.fail(id + ": the JWL EOS is used with 'hiex' only");
    }

    mpm_problem* p = (mpm_problem*)calloc(1, sizeof(mpm_problem));
    const int n = d.nb_particle;
    p->nb_particle = n;
    p->nb_body = d.nb_body;
    p->nb_mat = d.nb_mat;
    p->shape = d.sgmp ? MPM_SHAPE_SGMP : d.gimp ? MPM_SHAPE_GIMP : d.bspline ? MPM_SHAPE_BSPLINE : MPM_SHAPE_LINEAR;
    // the reference tests GIMP before Bspline in every pass (update_step.f90:365-373) and
    // SGMP before both; the flags are exclusive in practice
    // MPM3D.f90 keys the order on USF and the re-map on MUSL; a deck switching
    // on both USF and MUSL mixes the two orders and is rejected.
    if (d.usf && d.musl) d.fail("'usf on' and 'musl on' together are not a supported scheme");
    p->scheme = d.usf ? MPM_USF : (d.musl ? MPM_MUSL : MPM_USL);
    p->quasi_load = d.quasi_load;
    p->quasi_damp = d.quasi_damp;
    p->gravity = d.gravity;
    p->update_v_by_fp = d.uvfp;
    for (int i = 0; i < 6; i++) p->fix[i] = d.fix[i];
    p->dcell = d.dcell;
    p->cutoff = d.cutoff;
    p->dtscale = d.dtscale;
    p->endtime = d.endtime;
    p->quasi_load_time = d.qlt;
    p->damp_factor = d.damp;
    p->bq1 = d.bq1;
    p->bq2 = d.bq2;
    p->dr_damping = d.dr_damping;
    p->smooth_stress = d.smooth_stress;
    p->nb_component = d.nb_component;
    p->tlmpm = d.tlmpm;
    if (d.tlmpm) {
      // TLConstitution stops for every model but the neo-Hookean one
      // (Constitution.f90:487-497)
      for (int m = 0; m < d.nb_mat; m++)
        if (d.mats[m].type != MPM_MAT_NEOH)
          d.fail("total-Lagrangian MPM supports the neo-Hookean material only (material " + std::to_string(m + 1) + ")");
    }
    p->contact = d.contact;
    p->fricfa = d.fricfa;
    p->normbody = d.normbody;
    if (d.contact) {
      // the reference leaves these cases undefined: SGMP's cell-centre passes
      // share one auxiliary grid between the components (comID keeps the
      // last body's value there), Lagr_NodContact reads components 1 and 2,
      // and another normbody leaves the normal unset
      if (d.sgmp) d.fail("contact with SGMP is not supported");
      if (d.nb_component < 2) d.fail("contact needs nbco >= 2");
      if (d.normbody < 0 || d.normbody > 2) d.fail("contact normal (normbody) must be 0, 1 or 2");
      for (int b = 0; b < d.nb_body; b++)
        if (d.bodies[b].com < 1 || d.bodies[b].com > d.nb_component)
          d.fail("body " + std::to_string(b + 1) + ": component outside 1..nbco");
    }
    p->mats = (mpm_material*)calloc(d.nb_mat, sizeof(mpm_material));
    p->bodies = (mpm_body*)calloc(d.nb_body, sizeof(mpm_body));
    for (int m = 0; m < d.nb_mat; m++) p->mats[m] = d.mats[m];
    for (int b = 0; b < d.nb_body; b++) p->bodies[b] = d.bodies[b];
    p->pstate = (double*)calloc((size_t)PF_COUNT * n, sizeof(double));
    p->pbody = (int32_t*)calloc(n, sizeof(int32_t));
    double* ps = p->pstate;
    for (int i = 0; i < n; i++) {
      for (int c = 0; c < 3; c++) {
        ps[(size_t)(PF_XO + c) * n + i] = d.xp[c][i];   // Xo = Xp (DataIn.f90:859-863)
        ps[(size_t)(PF_X + c) * n + i] = d.xp[c][i];
        ps[(size_t)(PF_VX + c) * n + i] = d.vxp[c][i];
        ps[(size_t)(PF_FX + c) * n + i] = d.fxp[c][i];
      }
      ps[(size_t)PF_MASS * n + i] = d.mass[i];
    }
    // Initial (DataIn.f90:1211-1267)
    for (int b = 0; b < d.nb_body; b++) {
      const mpm_body& bd = p->bodies[b];
      mpm_material& mt = p->mats[bd.mat - 1];
      for (int i = bd.par_begin; i < bd.par_end; i++) {
        p->pbody[i] = b;
        double vol = ps[(size_t)PF_MASS * n + i] / mt.density;
        ps[(size_t)PF_VOL * n + i] = vol;
        ps[(size_t)PF_VOL0 * n + i] = vol;
        ps[(size_t)PF_SIGY * n + i] = mt.yield0;
        ps[(size_t)PF_IE * n + i] = mt.ceos[9] * vol;
        if (mt.type == MPM_MAT_ELASTIC || mt.type == MPM_MAT_NEOH) {
          double E = mt.young, nu = mt.poisson, ro = mt.density;
          ps[(size_t)PF_CP * n + i] = std::sqrt(E * (1 - nu) / (1 + nu) / (1 - 2 * nu) / ro);
        }
        if (mt.type == MPM_MAT_NULL) {
          ps[(size_t)PF_CP * n + i] = mt.wavespd;
          mt.ceos[0] = mt.wavespd;
        }
        // lighting time: the nearest detonation point's distance over D,
        // through a default REAL (DataIn.f90:1252-1260)
        double lt = (double)10e6f;
        for (int j = 0; j < d.n_deto; j++) {
          const double dx = ps[(size_t)PF_X * n + i] - d.deto[j][0], dy = ps[(size_t)PF_Y * n + i] - d.deto[j][1],
                       dz = ps[(size_t)PF_Z * n + i] - d.deto[j][2];
          const float temp = (float)(std::sqrt(dx * dx + dy * dy + dz * dz) / mt.det_speed);
          lt = (double)temp < lt ? (double)temp : lt;
        }
        ps[(size_t)PF_LT * n + i] = lt;
        ps[(size_t)PF_TEMP * n + i] = (double)293.0f;   // InitParticle (Particle.f90:131)
      }
    }
    // SetGridData (Grid.f90:238-263)
    double spx = d.span[1] - d.span[0], spy = d.span[3] - d.span[2], spz = d.span[5] - d.span[4];
    if (spx <= 0 || spy <= 0 || spz <= 0) d.fail("SPX/SPY/SPZ must be defined");
    p->ncx = (int)(spx / d.dcell + 0.5);
    p->ncy = (int)(spy / d.dcell + 0.5);
    p->ncz = (int)(spz / d.dcell + 0.5);
    p->span[0] = d.span[0];
    p->span[2] = d.span[2];
    p->span[4] = d.span[4];
    p->span[1] = d.span[0] + (double)p->ncx * d.dcell;
    p->span[3] = d.span[2] + (double)p->ncy * d.dcell;
    p->span[5] = d.span[4] + (double)p->ncz * d.dcell;
    // SetDT (DataIn.f90:1269-1297)
    double dt = (double)1.0e6f;
    for (int m = 0; m < d.nb_mat; m++) {
      const mpm_material& mt = p->mats[m];
      double E = mt.young, nu = mt.poisson, ro = mt.density;
      double cp = std::sqrt(E * (1 - nu) / (1 + nu) / (1 - 2 * nu) / ro);
      cp = std::fmax(cp, mt.det_speed);
      cp = std::fmax(cp, mt.wavespd);
      dt = std::fmin(dt, d.dcell / cp);
    }
    p->dt = dt * d.dtscale;
    *out = p;
    return MPM_OK;
  } catch (const DeckError& e) {
    if (d.tk.f) fclose(d.tk.f);
    if (errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", e.msg.c_str());
    return MPM_ERR_DECK;
  }
}

extern "C" void mpm_problem_free(mpm_problem* p) {
  if (!p) return;
  free(p->mats);
  free(p->bodies);
  free(p->pstate);
  free(p->pbody);
  free(p);
}

extern "C" int mpm_deck_read_outputs(const char* path, mpm_outputs* out, char* errbuf, int errlen) {
  if (!out || !path) return MPM_ERR_ARG;
  Deck d;
  d.tk.f = fopen(path, "r");
  if (!d.tk.f) {
    if (errbuf && errlen > 0) snprintf(errbuf, errlen, "cannot open deck '%s'", path);
    return MPM_ERR_IO;
  }
  try {
    d.parse();
  } catch (const DeckError& e) {
    fclose(d.tk.f);
    if (errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", e.msg.c_str());
    return MPM_ERR_DECK;
  }
  fclose(d.tk.f);
  *out = d.outs;
  return MPM_OK;
}
