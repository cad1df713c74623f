// accp_run.cpp -- the Accel+ command-file driver on the GPU backend.
//
//   accp-run [--precision N] [--specialized] [--parse-only] command_file
//
// Reads the reference's command files (src/process.cpp) and runs their
// Lattice, Optics, Track and DAMA blocks through libaccp: lattice files in
// the reference grammar (src/lib/accel/accel.cpp:67-133 and each create()),
// the same keys, defaults and printed blocks ($$$...), tracking, optics and
// DA on the GPU.  Numbers print with 6 significant digits like the
// reference's cout unless --precision asks for more.  --specialized runs
// Do_Track on the kernel compiled at run time for the lattice's structure
// (ACCP_STAGE_SPECIALIZED: about half a second of compilation, then a
// faster turn; long runs gain).
// --parse-only reads the command file and the lattice and prints the element
// count without touching the GPU.
//
// Commands handled (the path's callers in the reference):
//   Lattice: file_name=F[, type=ERING]; set_flags: radiation=, cavity=, wake=;
//            export_element: type=RING|T[, famname=F]; apply_setpt=file;
//            modify_element: type=, famname=, index=, field=, value=|delta=;
//            correct_tune / correct_chrom: type=, famname=(F1 F2) [a, b]
//            [REPEAT] [METHOD=FD|ADJOINT]; set_max_num_threads=N (accepted)
//   Optics:  print_tune / print_chrom [dpp=, delta=]; print_2nd_chrom;
//            print_matrix: type=RING|T, famname=, index=n|[i j ...], line_style=;
//            print_optics: type=, famname=, index=n|[i j ...]
//   Track:   Distribution: Zero NPar= | Load f[, format=plain|accel-export] |
//            Coordinate r=(...;...) | Gaussian NPar= sigx=|emit_x=[,betx=] ... |
//            Linear X=a:d:b
//            [shift=(...)] [print]; Monitor: interval=, fraction=,
//            printflag=Coordinate|Stat1|Stat2; Do_Track: NTurn=,
//            radiation=, RFCavity=, PrintFinal=No|Yes|Full [, seed=];
//            Transport: type=, famname=, index=n|[i j ...]
//   DAMA:    TRACK_DA: xmin=, xmax=, ymax=, nray=, nstep=, nturn=
//            [, print=loss_turn] [, weightminusside=true]; apply_setpt=file;
//            TRACK_MA: type=, [famname=,] nmoni=, dppmin=, dppmax=, nstep=, nturn=;
//            TRACK_DNUDJ: xmax=, ymax=, nturn=
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/accp.h"

namespace {

using std::string;

int g_prec = 6;
bool g_specialized = false;  // --specialized: Do_Track on the specialized kernel

// Tracking options of Do_Track: the launch shape that suits the beam (64-lane
// workgroups for beams that leave the GPU partly idle, 256 otherwise).
accp_track_opts track_opts(int np) {
  accp_track_opts o{};
  if (g_specialized) {
    o.tuning.staging = ACCP_STAGE_SPECIALIZED;
    o.tuning.block = np >= 65536 ? 256 : 64;
    o.tuning.ppl = 1;
  }
  return o;
}

// ---- text helpers (process.cpp:109-160) ----------------------------------------
string trim(const string &s) {
  size_t a = s.find_first_not_of(" \t\r\n\f\v");
  if (a == string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n\f\v");
  return s.substr(a, b - a + 1);
}
string upper(string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}
bool ieq(const string &a, const string &b) { return upper(trim(a)) == upper(trim(b)); }

void strip_comment(string &s) {
  size_t p = s.find_first_of("#%");
  if (p != string::npos) s.erase(p);
}

// One logical line: comments removed, '\' continues a line.
bool get_logical_line(std::istream &in, string &line) {
  if (!std::getline(in, line)) return false;
  strip_comment(line);
  line = trim(line);
  while (!line.empty() && line.back() == '\\') {
    line.back() = ' ';
    string next;
    if (!std::getline(in, next)) break;
    strip_comment(next);
    line += trim(next);
  }
  return true;
}

// The next line whose upper-cased text starts with one of keys; -1 at
// STOP/RETURN or end of file (process.cpp:136-160).
int look_for(std::istream &in, string &line, const std::vector<string> &keys) {
  string raw;
  while (get_logical_line(in, raw)) {
    line = trim(raw);
    const string up = upper(line);
    if (up == "STOP" || up == "RETURN") return -1;
    for (size_t i = 0; i < keys.size(); ++i)
      if (up.compare(0, keys[i].size(), keys[i]) == 0) return (int)i;
  }
  line.clear();
  return -1;
}

// "key=value" pairs after the first '=' or ':', commas as blanks.
std::vector<std::pair<string, string>> key_values(string s) {
  size_t p = s.find_first_of("=:");
  s = p == string::npos ? "" : s.substr(p + 1);
  std::replace(s.begin(), s.end(), ',', ' ');
  std::vector<std::pair<string, string>> kv;
  s = trim(s);
  while (!s.empty()) {
    p = s.find('=');
    if (p == string::npos) {
      kv.emplace_back(upper(trim(s)), "");
      break;
    }
    string key = upper(trim(s.substr(0, p)));
    s = trim(s.substr(p + 1));
    size_t q = s.find_first_of(" \t");
    kv.emplace_back(key, s.substr(0, q));
    s = q == string::npos ? "" : trim(s.substr(q + 1));
  }
  return kv;
}

string fmt(double v) {
  std::ostringstream o;
  o << std::setprecision(g_prec) << v;
  return o.str();
}
string fmt_sci(double v) {
  std::ostringstream o;
  o << std::scientific << std::setprecision(g_prec) << v;
  return o.str();
}

void check(int rc, const char *what) {
  if (rc != ACCP_OK) throw std::runtime_error(string(what) + ": " + accp_last_error());
}

// ---- lattice file (accel.cpp:67-133 and the create() of each type) ------------
accp_elem new_elem(int type, const string &fam) {
  accp_elem e;
  std::memset(&e, 0, sizeof e);
  e.type = type;
  e.nslice = 10;      // accel.h:195
  e.energy = 3000.0;  // BendMPole default energy [MeV]
  e.on = 1;
  std::strncpy(e.famname, fam.c_str(), ACCP_NAME_LEN - 1);
  return e;
}

// WigTable rows: n numbers from one line of the file.  The reference appends
// each line to one stream and reads across the joins (accel.cpp:1298-1335),
// so a line that does not end in a blank fuses with the next one; the values
// are read line by line here, which is the same for files written by
// serialize (every value followed by a tab).
void read_row(std::istream &in, double *out, int n, const string &fam) {
  string row;
  if (!std::getline(in, row)) throw std::runtime_error("WigTable " + fam + ": table ends early");
  std::istringstream rs(row);
  for (int i = 0; i < n; ++i)
    if (!(rs >> out[i])) throw std::runtime_error("WigTable " + fam + ": short table row");
}

std::vector<accp_elem> read_lattice(const string &path, std::deque<std::vector<double>> &tables) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open lattice file " + path);
  std::vector<accp_elem> v;
  string fam;
  while (in >> fam) {
    if (fam == "RETURN") break;
    if (fam[0] == '!') {  // comment line
      string skip;
      std::getline(in, skip);
      continue;
    }
    if (fam.back() == ':') fam.pop_back();
    else {
      string skip;
      if (!std::getline(in, skip, ':')) break;
    }
    string type, rest;
    if (!(in >> type)) break;
    std::getline(in, rest);
    std::istringstream ps(rest);
    accp_elem e;
    if (type == "Drift") {
      e = new_elem(ACCP_DRIFT, fam);
      ps >> e.len;
    } else if (type == "Quad") {
      e = new_elem(ACCP_QUAD, fam);
      ps >> e.len >> e.k;
    } else if (type == "Corrector") {
      e = new_elem(ACCP_CORRECTOR, fam);
      ps >> e.len >> e.kickx >> e.kicky;
    } else if (type == "Aperture") {
      e = new_elem(ACCP_APERTURE, fam);
      ps >> e.len >> e.xlim[0] >> e.xlim[1] >> e.ylim[0] >> e.ylim[1];
    } else if (type == "Monitor") {
      e = new_elem(ACCP_MONITOR, fam);
      ps >> e.len;
    } else if (type == "Bend") {
      e = new_elem(ACCP_BEND, fam);
      ps >> e.len >> e.theta >> e.e1 >> e.e2 >> e.k;
    } else if (type == "BendMPole") {
      e = new_elem(ACCP_BENDMPOLE, fam);
      ps >> e.len >> e.theta >> e.e1 >> e.e2 >> e.max_npole;
      if (e.max_npole < 1 || e.max_npole > ACCP_MAX_MPOLE)
        throw std::runtime_error("BendMPole " + fam + ": MaxNPole out of range");
      for (int i = 0; i < e.max_npole; ++i) in >> e.polyB[i];
      std::getline(in, rest);
    } else if (type == "StrMPole") {
      e = new_elem(ACCP_STRMPOLE, fam);
      ps >> e.len >> e.max_npole;
      if (e.max_npole < 1 || e.max_npole > ACCP_MAX_MPOLE)
        throw std::runtime_error("StrMPole " + fam + ": MaxNPole out of range");
      for (int i = 0; i < e.max_npole; ++i) in >> e.polyB[i];
      std::getline(in, rest);
      for (int i = 0; i < e.max_npole; ++i) in >> e.polyA[i];
      std::getline(in, rest);
    } else if (type == "Cavity") {
      e = new_elem(ACCP_CAVITY, fam);
      ps >> e.len >> e.volt >> e.freq >> e.phis >> e.energy;
    } else if (type == "Matrix") {  // accel.cpp:1673-1687: L, then six rows
      e = new_elem(ACCP_MATRIX, fam);
      ps >> e.len;
      for (int i = 0; i < 6; ++i) {
        string row;
        std::getline(in, row);
        std::istringstream rs(row);
        for (int j = 0; j < 6; ++j) rs >> e.mat[6 * i + j];
      }
    } else if (type == "SyncRad") {  // accel.cpp:1743-1753: L E U0 Jx Jy / ex ey sigdpp bx by
      e = new_elem(ACCP_SYNCRAD, fam);
      ps >> e.len >> e.energy >> e.u0 >> e.jx >> e.jy;
      string row;
      std::getline(in, row);
      std::istringstream(row) >> e.emitx >> e.emity >> e.sigdpp >> e.betax >> e.betay;
    } else if (type == "RFMult") {  // accel.cpp:1058-1068: L f phis E sigphi sigdVV MaxNPole / PolyB / PolyA
      e = new_elem(ACCP_RFMULT, fam);
      ps >> e.len >> e.freq >> e.phis >> e.energy >> e.sigphi >> e.sigdvv >> e.max_npole;
      // The reference parses the noise sigmas, then builds the element with
      // a constructor that clears them (accel.cpp:1065-1071, accel.h:376):
      // an RFMult read from a lattice file has no phase noise.
      e.sigphi = 0;
      e.sigdvv = 0;
      if (e.max_npole < 0 || e.max_npole > ACCP_MAX_MPOLE)
        throw std::runtime_error("RFMult " + fam + ": MaxNPole out of range");
      for (int i = 0; i < e.max_npole; ++i) in >> e.polyB[i];
      std::getline(in, rest);
      for (int i = 0; i < e.max_npole; ++i) in >> e.polyA[i];
      std::getline(in, rest);
    } else if (type == "IdealWig") {  // accel.cpp:1449-1456: L nslice lambda h type
      e = new_elem(ACCP_IDEALWIG, fam);
      ps >> e.len >> e.nslice >> e.wig_lambda >> e.wig_h >> e.wig_type;
    } else if (type == "TransWake") {  // accel.cpp:1546-1554: L E Ne Wx fRx Qx Wy fRy Qy
      e = new_elem(ACCP_TRANSWAKE, fam);
      ps >> e.len >> e.energy >> e.wake_charge >> e.tw_w[0] >> e.tw_fr[0] >> e.tw_q[0] >> e.tw_w[1] >> e.tw_fr[1] >>
          e.tw_q[1];
    } else if (type == "WigTable") {  // accel.cpp:1298-1335: L nslice mx my / xl / yl / xkick rows / blank / ykick rows
      e = new_elem(ACCP_WIGTABLE, fam);
      ps >> e.len >> e.nslice >> e.wt_mx >> e.wt_my;
      const int mx = e.wt_mx, my = e.wt_my;
      if (mx < 2 || mx > ACCP_WIGTAB_MAX_X || my < 2 || my > ACCP_WIGTAB_MAX_Y)
        throw std::runtime_error("WigTable " + fam + ": mx must be 2..101 and my 2..21");
      tables.emplace_back((size_t)mx + my + 2 * (size_t)mx * my);
      double *t = tables.back().data();
      read_row(in, t, mx, fam);
      read_row(in, t + mx, my, fam);
      for (int i = 0; i < my; ++i) read_row(in, t + mx + my + (size_t)i * mx, mx, fam);
      std::getline(in, rest);  // the blank line between the two tables
      for (int i = 0; i < my; ++i) read_row(in, t + mx + my + (size_t)(my + i) * mx, mx, fam);
      e.wt_table = t;
    } else if (type == "CrabCav") {  // accel.cpp:937-943
      e = new_elem(ACCP_CRABCAV, fam);
      ps >> e.len >> e.devoltx >> e.devolty >> e.freq >> e.phis >> e.energy >> e.sigphi >> e.sigdvv;
    } else if (type == "LongWake") {  // new element: L Rs Q fr charge zmin zmax nbins E
      e = new_elem(ACCP_LONGWAKE, fam);
      ps >> e.len >> e.wake_rs >> e.wake_q >> e.wake_fr >> e.wake_charge >> e.wake_zmin >> e.wake_zmax >>
          e.nbins >> e.energy;
    } else {
      throw std::runtime_error("element " + fam + ": type " + type + " is not carried by this backend");
    }
    v.push_back(e);
  }
  if (v.empty()) throw std::runtime_error("no element read from " + path);
  return v;
}

// ---- the line: host records plus the device table ------------------------------
struct Line {
  std::vector<accp_elem> e;
  std::deque<std::vector<double>> tables;  // WigTable data read from the lattice file
  accp_lattice *lat = nullptr;
  bool dirty = true;
  bool parse_only = false;

  ~Line() { accp_lattice_free(lat); }
  accp_lattice *device() {
    if (parse_only) throw std::runtime_error("--parse-only: no GPU work");
    if (!lat) check(accp_lattice_build(e.data(), (int)e.size(), &lat), "accp_lattice_build");
    else if (dirty) check(accp_lattice_update(lat, e.data(), (int)e.size()), "accp_lattice_update");
    dirty = false;
    return lat;
  }
  void pull() {  // after a library call changed the lattice
    check(accp_lattice_get(lat, e.data(), (int)e.size()), "accp_lattice_get");
    dirty = false;
  }
  static const char *type_name(int t) {
    switch (t) {
      case ACCP_DRIFT: return "Drift";
      case ACCP_QUAD: return "Quad";
      case ACCP_CORRECTOR: return "Corrector";
      case ACCP_APERTURE: return "Aperture";
      case ACCP_MONITOR: return "Monitor";
      case ACCP_BEND: return "Bend";
      case ACCP_BENDMPOLE: return "BendMPole";
      case ACCP_STRMPOLE: return "StrMPole";
      case ACCP_CAVITY: return "Cavity";
      case ACCP_LONGWAKE: return "LongWake";
      case ACCP_MATRIX: return "Matrix";
      case ACCP_SYNCRAD: return "SyncRad";
      case ACCP_CRABCAV: return "CrabCav";
      case ACCP_RFMULT: return "RFMult";
      case ACCP_IDEALWIG: return "IdealWig";
      case ACCP_TRANSWAKE: return "TransWake";
      case ACCP_WIGTABLE: return "WigTable";
    }
    return "?";
  }
  // ALine::findelements (aline.cpp:520-555); index = 1 for the first match,
  // 0 for all of them.
  std::vector<int32_t> find(const string &type, const string &fam, unsigned index = 0) const {
    std::vector<int32_t> out;
    unsigned cnt = 0;
    for (size_t i = 0; i < e.size(); ++i)
      if (type == type_name(e[i].type) && (fam.empty() || fam == e[i].famname)) {
        ++cnt;
        if (index == 0) out.push_back((int32_t)i);
        else if (index == cnt) return {(int32_t)i};
      }
    return out;
  }
  // ALine::set_rad_onoff / set_cavity_onoff / set_wake_onoff (aline.cpp:258-293, :429-463)
  int set_rad(int flag) {
    int n = 0;
    for (auto &x : e)
      if (x.type == ACCP_BENDMPOLE || x.type == ACCP_IDEALWIG) x.rad = flag <= ACCP_RAD_QE ? flag : ACCP_RAD_OFF, ++n;
      else if (x.type == ACCP_SYNCRAD) x.rad = flag >= ACCP_RAD_LUMP ? flag : ACCP_RAD_OFF, ++n;
    dirty = true;
    return n;
  }
  int set_on(int type, int on) {
    int n = 0;
    for (auto &x : e)
      if (x.type == type || (type == ACCP_LONGWAKE && x.type == ACCP_TRANSWAKE)) x.on = on, ++n;
    dirty = true;
    return n;
  }
  double length() const {
    double s = 0;
    for (const auto &x : e) s += x.len;
    return s;
  }
};

// AccelElement::setfield/getfield names (accel.cpp) for modify_element.
double *field_ptr(accp_elem &x, const string &f, unsigned index) {
  if (f == "Length") return &x.len;
  if (f == "K" && (x.type == ACCP_QUAD || x.type == ACCP_BEND)) return &x.k;
  if (f == "BendingAngle") return &x.theta;
  if (f == "EntranceAngle") return &x.e1;
  if (f == "ExitAngle") return &x.e2;
  if (f == "energy" || f == "Energy") return &x.energy;
  if (f == "Voltage") return &x.volt;
  if (f == "Frequency") return &x.freq;
  if (f == "Phis") return &x.phis;
  if (f == "KickX") return &x.kickx;
  if (f == "KickY") return &x.kicky;
  if ((f == "PolynomB" || f == "PolynomA") && index < ACCP_MAX_MPOLE &&
      (x.type == ACCP_BENDMPOLE || x.type == ACCP_STRMPOLE)) {
    if (f == "PolynomA" && x.type == ACCP_BENDMPOLE) return nullptr;  // accel.cpp:630-634
    return f == "PolynomB" ? &x.polyB[index] : &x.polyA[index];
  }
  return nullptr;
}

// AccelElement::serialize of each type (accel.cpp:143-858): the lattice-file
// grammar, at 17 significant digits so that a file written here reads back
// to the same doubles.
string lossless(double v) {
  std::ostringstream o;
  o << std::setprecision(17) << v;
  return o.str();
}

void serialize(const accp_elem &x, std::ostream &os) {
  const string head = string(x.famname) + "  : \t" + Line::type_name(x.type) + "\t" + lossless(x.len);
  auto vals = [&](std::initializer_list<double> v) {
    string t;
    for (double d : v) t += "\t" + lossless(d);
    return t;
  };
  switch (x.type) {
    case ACCP_DRIFT: os << head << "\t" << std::endl; break;
    case ACCP_QUAD: os << head << vals({x.k}) << std::endl; break;
    case ACCP_CORRECTOR: os << head << vals({x.kickx, x.kicky}) << std::endl; break;
    case ACCP_APERTURE: os << head << vals({x.xlim[0], x.xlim[1], x.ylim[0], x.ylim[1]}) << std::endl; break;
    case ACCP_MONITOR: os << head << "\t1\t1\t0" << std::endl; break;
    case ACCP_BEND: os << head << vals({x.theta, x.e1, x.e2, x.k}) << std::endl; break;
    case ACCP_BENDMPOLE:
      os << head << vals({x.theta, x.e1, x.e2}) << "\t" << x.max_npole << std::endl << "\t\t\t";
      for (int i = 0; i < x.max_npole; ++i) os << lossless(x.polyB[i]) << "\t";
      os << std::endl;
      break;
    case ACCP_STRMPOLE:
      os << head << "\t" << x.max_npole << std::endl << "\t\t\t";
      for (int i = 0; i < x.max_npole; ++i) os << lossless(x.polyB[i]) << "\t";
      os << std::endl << "\t\t\t";
      for (int i = 0; i < x.max_npole; ++i) os << lossless(x.polyA[i]) << "\t";
      os << std::endl;
      break;
    case ACCP_CAVITY: os << head << vals({x.volt, x.freq, x.phis, x.energy}) << std::endl; break;
    case ACCP_MATRIX:
      os << head << "\t" << std::endl;
      for (int i = 0; i < 6; ++i) {
        os << "\t";
        for (int j = 0; j < 6; ++j) os << lossless(x.mat[6 * i + j]) << "\t";
        os << std::endl;
      }
      break;
    case ACCP_SYNCRAD:
      os << head << vals({x.energy, x.u0, x.jx, x.jy}) << "\t" << std::endl
         << vals({x.emitx, x.emity, x.sigdpp, x.betax, x.betay}) << "\t" << std::endl;
      break;
    case ACCP_RFMULT:
      os << head << vals({x.freq, x.phis, x.energy, x.sigphi, x.sigdvv}) << "\t" << x.max_npole << std::endl
         << "\t\t\t";
      for (int i = 0; i < x.max_npole; ++i) os << lossless(x.polyB[i]) << "\t";
      os << std::endl << "\t\t\t";
      for (int i = 0; i < x.max_npole; ++i) os << lossless(x.polyA[i]) << "\t";
      os << std::endl;
      break;
    case ACCP_IDEALWIG:
      os << head << "\t" << x.nslice << vals({x.wig_lambda, x.wig_h}) << "\t" << x.wig_type << std::endl;
      break;
    case ACCP_TRANSWAKE:
      os << head << vals({x.energy, x.wake_charge, x.tw_w[0], x.tw_fr[0], x.tw_q[0], x.tw_w[1], x.tw_fr[1], x.tw_q[1]})
         << std::endl;
      break;
    case ACCP_CRABCAV:
      os << head << vals({x.devoltx, x.devolty, x.freq, x.phis, x.energy, x.sigphi, x.sigdvv}) << std::endl;
      break;
    case ACCP_WIGTABLE: {  // AWigTable::serialize (accel.cpp:1246-1270)
      const int mx = x.wt_mx, my = x.wt_my;
      const double *t = x.wt_table;
      os << head << "\t" << x.nslice << "\t" << mx << "\t" << my << std::endl;
      auto row = [&](const double *v, int n) {
        for (int i = 0; i < n; ++i) os << lossless(v[i]) << "\t";
        os << std::endl;
      };
      row(t, mx);
      row(t + mx, my);
      for (int i = 0; i < my; ++i) row(t + mx + my + (size_t)i * mx, mx);
      os << std::endl;
      for (int i = 0; i < my; ++i) row(t + mx + my + (size_t)(my + i) * mx, mx);
      break;
    }
    case ACCP_LONGWAKE:
      os << head << vals({x.wake_rs, x.wake_q, x.wake_fr, x.wake_charge, x.wake_zmin, x.wake_zmax}) << "\t"
         << x.nbins << "\t" << lossless(x.energy) << std::endl;
      break;
  }
}

// type=..., famname=... of a command (process.cpp:2221-2268).
void type_famname(const string &line, string &type, string &fam) {
  for (const auto &kv : key_values(line)) {
    if (kv.first == "TYPE") type = kv.second;
    else if (kv.first == "FAMNAME") fam = kv.second;
  }
}

// EXPORT_ELEMENT (process.cpp:505-547)
void cmd_export(Line &L, const string &line) {
  string type, fam;
  type_famname(line, type, fam);
  std::cout << "$$$Export-Elements***********************" << std::endl;
  if (ieq(type, "RING") || ieq(type, "LINE")) {
    for (const auto &x : L.e) serialize(x, std::cout);
  } else {
    for (int32_t i : L.find(type, fam)) serialize(L.e[i], std::cout);
    std::cout << "******************************************" << std::endl;
  }
}

// apply_line_setpt (process.cpp:2273-2392): a "$FirstElement: type=, famname="
// line that must name the line's first element (zero-length Monitors
// before it are skipped), then "$SetValue: type=StrMPole, famname=F,
// PolyB[n]" (or PolyA[n]) lines, each followed by one value per family
// member.  Only StrMPole set points are applied, as in the reference.
void cmd_apply_setpt(Line &L, const string &line) {
  string file = trim(line.substr(line.find_first_of("=:") + 1));
  size_t e = file.find('=');
  if (e != string::npos) file = trim(file.substr(e + 1));
  std::cout << file << std::endl;
  std::ifstream in(file);
  if (!in) throw std::runtime_error("cannot open set-point file " + file);
  string sl;
  bool first_ok = false;
  while (get_logical_line(in, sl)) {
    if (!ieq(sl.substr(0, 13), "$FirstElement")) continue;
    string type, fam;
    type_famname(sl, type, fam);
    for (const auto &x : L.e) {
      if (type == Line::type_name(x.type) && fam == x.famname) {
        first_ok = true;
        break;
      }
      if (x.len > 0 || x.type != ACCP_MONITOR) break;
    }
    break;
  }
  if (!first_ok) {
    std::cout << "#first element check: 0" << std::endl;
    return;
  }
  std::cout << "#first element matches" << std::endl;
  while (get_logical_line(in, sl)) {
    if (!ieq(sl.substr(0, 9), "$SetValue")) continue;
    string type, fam;
    type_famname(sl, type, fam);
    const std::vector<int32_t> list = L.find(type, fam);
    std::vector<double> val(list.size());
    for (double &v : val) in >> v;
    if (type != "StrMPole") continue;
    const size_t p = sl.find("Poly");
    if (p == string::npos) continue;
    const bool b = sl.compare(p, 5, "PolyB") == 0;
    const size_t a = sl.find('[', p), z = sl.find(']', p);
    const int index = std::stoi(sl.substr(a + 1, z - a - 1));
    if (index < 0 || index >= ACCP_MAX_MPOLE) throw std::runtime_error("set point: multipole index out of range");
    for (size_t i = 0; i < list.size(); ++i) (b ? L.e[list[i]].polyB : L.e[list[i]].polyA)[index] = val[i];
    L.dirty = true;
  }
}

// ---- Lattice block (process.cpp:173-988) ------------------------------------------
void cmd_file_name(Line &L, const string &line) {
  string s = line;
  std::replace(s.begin(), s.end(), ',', ' ');
  size_t p = s.find_first_of("=:");
  s = trim(s.substr(p + 1));
  size_t q = s.find_first_of(" \t");
  const string file = s.substr(0, q);
  L.tables.clear();
  L.e = read_lattice(file, L.tables);
  if (L.lat) {
    accp_lattice_free(L.lat);
    L.lat = nullptr;
  }
  L.dirty = true;
  const string rest = q == string::npos ? "" : s.substr(q + 1);
  for (const auto &kv : key_values("x=" + rest))
    if (kv.first == "TYPE" && ieq(kv.second, "ERING") && !L.parse_only) {
      double phis = 0;
      check(accp_setrfphase(L.device(), &phis), "accp_setrfphase");
      L.pull();
      std::cout << "RF cavities phase set to: phi_s=" << fmt(phis) << " rad" << std::endl;
      L.set_rad(ACCP_RAD_OFF);
      std::cout << "radiation off" << std::endl;
    }
  const int ncav = L.set_on(ACCP_CAVITY, 0);
  if (ncav > 0) std::cout << ncav << " cavity(ies) turned off" << std::endl;
  const int nwake = L.set_on(ACCP_LONGWAKE, 0);
  if (nwake > 0) std::cout << nwake << " wake element(s) turned off" << std::endl;
  const int ncrab = L.set_on(ACCP_CRABCAV, 0);
  if (ncrab > 0) std::cout << ncrab << " crab cavity(ies) turned off" << std::endl;
}

void cmd_set_flags(Line &L, const string &line) {
  for (const auto &kv : key_values(line)) {
    const string &v = kv.second;
    if (kv.first == "RADIATION") {
      if (ieq(v, "off")) L.set_rad(ACCP_RAD_OFF), std::cout << "radiation turned off" << std::endl;
      else if (ieq(v, "rad") || ieq(v, "on"))
        L.set_rad(ACCP_RAD_DAMP), std::cout << "radiation turned on without quantum excitation" << std::endl;
      else if (ieq(v, "radQE"))
        L.set_rad(ACCP_RAD_QE), std::cout << "radiation turned on with quantum excitation" << std::endl;
    } else if (kv.first == "CAVITY" || kv.first == "WAKE" || kv.first == "CRABCAV") {
      const int type = kv.first == "CAVITY" ? ACCP_CAVITY : kv.first == "WAKE" ? ACCP_LONGWAKE : ACCP_CRABCAV;
      const char *what = kv.first == "CAVITY"   ? " cavity(ies) turned "
                         : kv.first == "WAKE" ? " wake element(s) turned "
                                              : " crab cavity(ies) turned ";
      if (ieq(v, "off") || ieq(v, "on")) {
        const int on = ieq(v, "on");
        const int n = L.set_on(type, on);
        if (n > 0) std::cout << n << what << (on ? "on" : "off") << std::endl;
      }
    }
  }
}

void cmd_modify_element(Line &L, const string &line) {
  string type, fam, field;
  unsigned index = 0;
  double value = 0;
  int mode = 0;  // 1 value, 2 delta
  for (const auto &kv : key_values(line)) {
    if (kv.first == "TYPE") type = kv.second;
    else if (kv.first == "FAMNAME") fam = kv.second;
    else if (kv.first == "INDEX") index = (unsigned)std::stoul(kv.second);
    else if (kv.first == "FIELD") field = kv.second;
    else if (kv.first == "VALUE") value = std::stod(kv.second), mode = 1;
    else if (kv.first == "DELTA") value = std::stod(kv.second), mode = 2;
  }
  // index selects the index-th family member (0: all) and, for PolynomB/A,
  // is also the multipole index, as the reference reads it (process.cpp:839-870)
  const std::vector<int32_t> list = L.find(type, fam, index);
  if (list.empty() || !mode) {
    if (list.empty()) std::cout << "No element is found: " << type << "\t" << fam << "\t" << index << std::endl;
    if (!mode) std::cout << "No value is given" << std::endl;
    return;
  }
  for (int32_t i : list) {
    accp_elem &x = L.e[i];
    if (field == "NSlice") {
      x.nslice = mode == 2 ? x.nslice + (int)value : (int)value;
    } else {
      double *f = field_ptr(x, field, index);
      if (!f) throw std::runtime_error("modify_element: field " + field + " not settable on " + type);
      *f = mode == 2 ? *f + value : value;
    }
  }
  L.dirty = true;
}

void cmd_correct(Line &L, const string &line, int quantity) {
  const bool tune = quantity == ACCP_TUNE;
  std::cout << (tune ? "$$$Correct Tunes***********************" : "$$$Correct chromaticities***********************")
            << std::endl;
  string s = line.substr(line.find_first_of("=:") + 1);
  std::replace(s.begin(), s.end(), ',', ' ');
  const string up = upper(s);
  string type;
  {
    size_t p = up.find("TYPE");
    size_t q = s.find('=', p);
    std::istringstream(s.substr(q + 1)) >> type;
  }
  std::cout << type << std::endl;
  string fam1, fam2;
  {
    size_t a = s.find('('), b = s.find(')');
    if (a == string::npos || b == string::npos) throw std::runtime_error("correct: famname=(F1 F2) expected");
    std::istringstream(s.substr(a + 1, b - a - 1)) >> fam1 >> fam2;
  }
  std::cout << fam1 << " " << fam2 << std::endl;
  double target[2] = {0, 0};
  {
    size_t a = s.find('['), b = s.find(']');
    if (a == string::npos || b == string::npos) throw std::runtime_error("correct: [target1 target2] expected");
    std::istringstream(s.substr(a + 1, b - a - 1)) >> target[0] >> target[1];
  }
  // Responses: by default the adjoint's exact derivatives of the getm66
  // tune; METHOD=FD asks for the reference's +1e-5 forward difference
  // (aline.cpp:1187-1242, 1310-1355), METHOD=ADJOINT insists on the adjoint.
  // Without METHOD, a lattice the reverse sweep cannot carry (noise,
  // quantum excitation, a wake) falls back to the forward difference.
  const bool want_fd = up.find("METHOD=FD") != string::npos;
  const bool want_adjoint = up.find("METHOD=ADJOINT") != string::npos;
  int method = want_fd ? ACCP_RESP_FD : ACCP_RESP_ADJOINT;
  std::cout << (tune ? "Target tunes: " : "Target chromaticities: ") << fmt(target[0]) << "\t" << fmt(target[1])
            << "\t" << std::endl;
  L.set_rad(ACCP_RAD_OFF);
  if (!tune) L.set_on(ACCP_CAVITY, 0);
  double v[2];
  if (tune) check(accp_calctune(L.device(), 0.0, ACCP_DELTA_DIFF, v), "accp_calctune");
  else check(accp_calcchrom(L.device(), 0.0, ACCP_DELTA_DIFF, v), "accp_calcchrom");
  std::cout << (tune ? "Initial tunes: " : "Initial chromaticities: ") << fmt(v[0]) << "\t" << fmt(v[1]) << "\t"
            << std::endl;
  const std::vector<int32_t> f1 = L.find(type, fam1), f2 = L.find(type, fam2);
  if (f1.empty() || f2.empty()) {
    std::cout << "error: no elements found " << f1.size() << "\t" << f2.size() << std::endl;
    return;
  }
  std::cout << f1.size() << "\t" << f2.size() << std::endl;
  const int steps = up.find("REPEAT") != string::npos ? 2 : 1;
  for (int k = 0; k < steps; ++k) {
    double t[2] = {target[0], target[1]};
    int rc = accp_correct(L.device(), quantity, method, f1.data(), (int)f1.size(), f2.data(), (int)f2.size(), t);
    if (rc == ACCP_E_LATTICE && method == ACCP_RESP_ADJOINT && !want_adjoint) {
      std::cout << "note: finite-difference responses (" << accp_last_error() << ")" << std::endl;
      method = ACCP_RESP_FD;
      t[0] = target[0];
      t[1] = target[1];
      rc = accp_correct(L.device(), quantity, method, f1.data(), (int)f1.size(), f2.data(), (int)f2.size(), t);
    }
    check(rc, "accp_correct");
    L.pull();
    std::cout << (tune ? "corrected tunes: " : "corrected chromaticities: ") << fmt(t[0]) << "\t" << fmt(t[1])
              << "\t" << std::endl;
  }
  std::cout << "*********************************" << std::endl;
}

// ALine::resetfirstelement (aline.cpp:557-605): rotate the line so that
// element i comes first.
void rotate_first(Line &L, int32_t i) {
  if (i <= 0) return;
  std::rotate(L.e.begin(), L.e.begin() + i, L.e.end());
  L.dirty = true;
}

// RESET_FIRST: type=, famname=[, index=] (process.cpp:254-340).  index picks
// the index-th member (the reference's findelements leaves the result of
// index > 1 unset; the intended member is taken).  With famname alone, each
// element of that name met while walking the (rotating) line moves the
// family's first member to the front, as the reference loops.
void cmd_reset_first(Line &L, const string &line) {
  string type, fam;
  unsigned index = 1;
  for (const auto &kv : key_values(line)) {
    if (kv.first == "TYPE") type = kv.second;
    else if (kv.first == "FAMNAME") fam = kv.second;
    else if (kv.first == "INDEX") index = (unsigned)std::stoul(kv.second);
  }
  if (!type.empty() && !fam.empty()) {
    std::cout << "First element moved to: type=" << type << ", famname=" << fam << ", index= " << index << std::endl;
    if (index == 0) std::cerr << "resetfirstelement: element index cannot be zero" << std::endl;
    const std::vector<int32_t> m = index ? L.find(type, fam, index) : std::vector<int32_t>{};
    if (m.empty())
      std::cout << "resetfirstelement: element not found (type =" << type << ", famname =" << fam
                << ", index =" << index << ")" << std::endl;
    else rotate_first(L, m[0]);
  } else if (type.empty() && !fam.empty()) {
    for (size_t i = 0; i < L.e.size(); ++i)
      if (fam == L.e[i].famname) {
        const string t = Line::type_name(L.e[i].type);
        std::cout << "First element moved to: type=" << t << ", famname=" << fam << ", index= " << 1 << std::endl;
        rotate_first(L, L.find(t, fam, 1)[0]);
      }
  }
}

// PRINT_ELEM_VALUE: type=[, famname=] then the names to print, for Quad
// (Index, Length, K) and StrMPole (Index, Length, PolyB[1..3], PolyA[1..3])
// (process.cpp:344-480).  Other types print an empty row per element.
void cmd_print_elem_value(Line &L, const string &line) {
  string s = line.substr(line.find_first_of("=:") + 1);
  std::replace(s.begin(), s.end(), ',', ' ');
  s = trim(s);
  // get_type_famname_str (process.cpp:2221-2267): two passes over
  // "key=value"; each pass drops up to the first blank after the '=' (or,
  // with no '=' left, the first word), so with type= alone the first name to
  // print is consumed, as in the reference
  string type, fam;
  int flag = -1;
  for (int pass = 0; pass < 2; ++pass) {
    const size_t e = s.find('=');
    const string key = trim(s.substr(0, e));
    s.erase(0, e + 1);  // e == npos: npos + 1 == 0, nothing erased
    s.erase(0, std::min(s.size(), s.find_first_not_of(" \t")));
    const size_t b = s.find_first_of(" \t");
    if (ieq(key, "type")) type = trim(s.substr(0, b)), flag = pass ? flag + 1 : 1;
    else if (ieq(key, "famname")) fam = trim(s.substr(0, b)), flag = pass ? flag + 10 : 10;
    s.erase(0, b + 1);
  }
  if (flag < 1) return;
  const std::vector<int32_t> list = L.find(type, flag == 11 ? fam : string());
  std::cout << "$$$Print-Element_Values***********************" << std::endl;
  for (int32_t i : list) {
    const accp_elem &x = L.e[i];
    if (type == "StrMPole") {
      if (s.find("Index") != string::npos) std::cout << i << "\t";
      if (s.find("Length") != string::npos) std::cout << fmt(x.len) << "\t";
      for (int k = 1; k <= 3; ++k)
        if (s.find("PolyB[" + std::to_string(k) + "]") != string::npos) std::cout << fmt(x.polyB[k]) << "\t";
      for (int k = 1; k <= 3; ++k)
        if (s.find("PolyA[" + std::to_string(k) + "]") != string::npos) std::cout << fmt(x.polyA[k]) << "\t";
    } else if (type == "Quad") {
      if (s.find("Index") != string::npos) std::cout << i << "\t";
      if (s.find("Length") != string::npos) std::cout << fmt(x.len) << "\t";
      if (s.find("K") != string::npos) std::cout << fmt(x.k) << "\t";
    }
    std::cout << std::endl;
  }
  std::cout << "******************************************" << std::endl;
}

// PRINT_LAT_SUMMARY (process.cpp:482-512): radiation off, tunes and
// chromaticities, then the 6D closed orbit (which leaves classical radiation
// and the cavities on, as ALine::getco6 does) and the ring-start Twiss
// functions of that state.
void cmd_lat_summary(Line &L) {
  std::cout << "$$$Lattice-Summary***********************" << std::endl;
  std::cout << "Total " << L.e.size() << " elements, total length = " << fmt(L.length()) << " m" << std::endl;
  std::cout << "First element is " << Line::type_name(L.e[0].type) << "." << L.e[0].famname << std::endl;
  L.set_rad(ACCP_RAD_OFF);
  double t[2], c[2], co[6], tw[7];
  check(accp_calctune(L.device(), 0.0, ACCP_DELTA_DIFF, t), "accp_calctune");
  std::cout << "Tunes: " << fmt(t[0]) << "\t" << fmt(t[1]) << "\t" << std::endl;
  check(accp_calcchrom(L.device(), 0.0, ACCP_DELTA_DIFF, c), "accp_calcchrom");
  std::cout << "Chromaticity: " << fmt(c[0]) << "\t" << fmt(c[1]) << "\t" << std::endl;
  check(accp_getco6(L.device(), co), "accp_getco6");
  L.pull();
  std::cout << "Closed orbit 6D " << std::endl;
  for (double v : co) std::cout << fmt(v) << "\t";
  std::cout << std::endl;
  check(accp_calctwiss(L.device(), nullptr, 0, 0.0, ACCP_DELTA_DIFF, tw), "accp_calctwiss");
  std::cout << "betax\tbetay\talphax\talphay" << std::endl;
  std::cout << fmt(tw[3]) << " \t" << fmt(tw[5]) << " \t" << fmt(tw[4]) << " \t" << fmt(tw[6]) << std::endl;
  std::cout << "******************************************" << std::endl;
}

void proc_lattice(std::istream &in, Line &L) {
  const std::vector<string> keys = {"FILE_NAME",      "CORRECT_TUNE", "CORRECT_CHROM", "SET_MAX_NUM_THREADS",
                                    "MODIFY_ELEMENT", "SET_FLAGS",    "RESET_FIRST",   "PRINT_ELEM_VALUE",
                                    "PRINT_LAT_SUMMARY", "EXPORT_ELEMENT", "APPLY_SETPT", "END"};
  string line;
  for (;;) {
    const int k = look_for(in, line, keys);
    if (k < 0 || keys[k] == "END") return;
    if (k != 0 && L.e.empty()) throw std::runtime_error("Lattice file name needs to be given first");
    if (keys[k] == "EXPORT_ELEMENT") {
      cmd_export(L, line);
      continue;
    }
    if (keys[k] == "APPLY_SETPT") {
      cmd_apply_setpt(L, line);
      continue;
    }
    if (keys[k] == "RESET_FIRST") {
      cmd_reset_first(L, line);
      continue;
    }
    if (keys[k] == "PRINT_ELEM_VALUE") {
      cmd_print_elem_value(L, line);
      continue;
    }
    if (L.parse_only && (k == 1 || k == 2 || k == 8)) continue;
    if (keys[k] == "PRINT_LAT_SUMMARY") {
      cmd_lat_summary(L);
      continue;
    }
    switch (k) {
      case 0: cmd_file_name(L, line); break;
      case 1: cmd_correct(L, line, ACCP_TUNE); break;
      case 2: cmd_correct(L, line, ACCP_CHROM); break;
      case 3: {
        int n = 1;
        std::istringstream(line.substr(line.find_first_of("=:") + 1)) >> n;
        std::cout << "Maximum thread number set to " << std::max(n, 1) << std::endl;  // host threads unused
        break;
      }
      case 4: cmd_modify_element(L, line); break;
      case 5: cmd_set_flags(L, line); break;
    }
  }
}

// get_elem_list (process.cpp:1050-1099): the elements a command selects.
// type empty: element `index`; ALL: every element; no famname: every
// element of the type, or its index-th (index > 0); "[i j ...]": the family
// members i, j, ... (1-based, kept only below the family size and written in
// place, as the reference does); else the family, or its index-th member
// (the reference's findelements leaves that result unset; the intended
// member is taken).
std::vector<int32_t> elem_list(const Line &L, const string &type, const string &fam, unsigned index,
                               const string &indices) {
  if (type.empty()) return {(int32_t)index};
  if (ieq(type, "ALL")) {
    std::vector<int32_t> v(L.e.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (int32_t)i;
    return v;
  }
  if (fam.empty()) {
    std::vector<int32_t> v = L.find(type, "");
    if (index > 0) v = index <= v.size() ? std::vector<int32_t>{v[index - 1]} : std::vector<int32_t>{};
    return v;
  }
  if (!indices.empty()) {
    std::vector<int32_t> v = L.find(type, fam);
    std::istringstream is(indices);
    int t, i = 0;
    while (is >> t)
      if (t > 0 && t < (int)v.size()) v[i++] = v[t - 1];
    v.resize(i);
    return v;
  }
  return L.find(type, fam, index);
}

// print_m66_to_line (process.cpp:1397-1425): "oneline" (36 values on one
// line), "4by4" (the transverse block) or "6by6"; every value followed by a
// tab.
void print_m66(const double *m, const string &style) {
  if (ieq(style, "oneline")) {
    for (int i = 0; i < 36; ++i) std::cout << fmt(m[i]) << "\t";
    std::cout << std::endl;
    return;
  }
  const int n = ieq(style, "4by4") ? 4 : ieq(style, "6by6") ? 6 : 0;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) std::cout << fmt(m[6 * i + j]) << "\t";
    std::cout << std::endl;
  }
}

// "index=[i j ...]" of a command: the list inside the brackets (commas as
// blanks), or "" when index is a single number or absent.
string index_list(const string &line) {
  const string up = upper(line);
  size_t p = up.find("INDEX");
  if (p == string::npos) return "";
  p = line.find('=', p);
  if (p == string::npos) return "";
  size_t a = line.find_first_not_of(" \t", p + 1);
  if (a == string::npos || line[a] != '[') return "";
  size_t b = line.find(']', a);
  string v = line.substr(a + 1, b == string::npos ? string::npos : b - a - 1);
  std::replace(v.begin(), v.end(), ',', ' ');
  return trim(v);
}

// The line with its bracketed index list replaced by 0 (so key=value
// parsing sees "index=0"; the list itself comes from index_list).
string strip_index_list(const string &line) {
  size_t a = line.find('['), b = line.find(']');
  if (a == string::npos || b == string::npos || b < a) return line;
  return line.substr(0, a) + "0" + line.substr(b + 1);
}

// ---- Optics block (process.cpp:1107-1395): tunes, chromaticities, one-turn matrix
void proc_optics(std::istream &in, Line &L) {
  const std::vector<string> keys = {"PRINT_MATRIX", "PRINT_TUNE", "PRINT_2ND_CHROM", "PRINT_CHROM",
                                    "PRINT_OPTICS", "END"};
  string line;
  for (;;) {
    const int k = look_for(in, line, keys);
    if (k < 0 || keys[k] == "END") return;
    double dpp = 0, delta = ACCP_DELTA_DIFF;
    string type, fam;
    unsigned index = 0;
    string indices, style = "6by6";
    const string bracketed = index_list(line);
    for (const auto &kv : key_values(bracketed.empty() ? line : strip_index_list(line))) {
      if (kv.first == "DPP") dpp = std::stod(kv.second);
      else if (kv.first == "DELTA") delta = std::stod(kv.second);
      else if (kv.first == "TYPE") type = kv.second;
      else if (kv.first == "FAMNAME") fam = kv.second;
      else if (kv.first == "INDEX") index = (unsigned)std::stoul(kv.second);
      else if (kv.first == "LINE_STYLE") style = kv.second;
    }
    indices = bracketed;
    if (L.parse_only) continue;
    if (k == 4) {
      // PRINT_OPTICS (process.cpp:1122-1226): Twiss at the selected elements
      std::vector<int32_t> list = elem_list(L, type, fam, index, indices);
      std::cout << "nlist=" << list.size() << std::endl;
      if (list.empty()) continue;
      std::vector<double> tw(7 * list.size());
      check(accp_calctwiss(L.device(), list.data(), (int)list.size(), dpp, delta, tw.data()), "accp_calctwiss");
      std::cout << "$$$Optics-Twiss, dp=" << fmt(dpp) << "*************************** " << list.size() << " points"
                << std::endl;
      std::cout << "index\ts (m)\tbetx (m)\tbety (m)\talfx\talfy\tpsix\tpsiy\n";
      for (size_t i = 0; i < list.size(); ++i) {
        const double *t = &tw[7 * i];
        std::cout << list[i] << "\t" << fmt(t[0]) << "\t" << fmt(t[3]) << "\t" << fmt(t[5]) << "\t" << fmt(t[4])
                  << "\t" << fmt(t[6]) << "\t" << fmt(t[1]) << "\t" << fmt(t[2]) << "\t" << std::endl;
      }
      continue;
    }
    L.set_rad(ACCP_RAD_OFF);
    if (k == 2) {
      // PRINT_2ND_CHROM (process.cpp:1380-1388)
      double c[4];
      check(accp_calcchrom2(L.device(), c), "accp_calcchrom2");
      std::cout << "Chromaticity: [" << fmt(c[0]) << "\t" << fmt(c[1]) << "]" << std::endl;
      std::cout << "2nd-order Chromaticity: [" << fmt(c[2]) << "\t" << fmt(c[3]) << "], " << std::endl;
      continue;
    }
    if (k == 0) {
      // PRINT_MATRIX (process.cpp:1228-1325): the one-turn matrix, or the
      // matrices from the ring start to each selected element
      double m[36];
      if (ieq(type, "RING") || ieq(type, "ONETURN")) {
        check(accp_getm66(L.device(), 0, -1, dpp, delta, m), "accp_getm66");
        std::cout << "$$$OPTICS-MATRIX*******************ONE-TURN" << std::endl;
        std::cout << "type=RING at " << Line::type_name(L.e[0].type) << ", famname=" << L.e[0].famname
                  << ", ring index=0, spos=" << fmt(L.length()) << std::endl;
        print_m66(m, style);
        std::cout << "****************************" << std::endl;
      } else {
        const std::vector<int32_t> list = elem_list(L, type, fam, index, indices);
        if (list.empty()) {
          std::cout << "#no element found for matrix print out, type=" << type << ", famname=" << fam << ",index"
                    << index << std::endl;
          continue;
        }
        std::cout << "$$$OPTICS-MATRIX, dpp=" << fmt(dpp) << "******************* " << list.size() << " points"
                  << std::endl;
        for (int32_t e : list) {
          if (e < 0 || e > (int)L.e.size()) throw std::runtime_error("PRINT_MATRIX: ring index out of range");
          check(accp_getm66(L.device(), 0, e, dpp, delta, m), "accp_getm66");
          double spos = 0;
          for (int i = 0; i < e; ++i) spos += L.e[i].len;
          std::cout << "type=" << type << ", famname=" << fam << ", ring index=" << e << ", spos=" << fmt(spos)
                    << std::endl;
          print_m66(m, style);
        }
        std::cout << "**********************" << std::endl;
      }
    } else if (k == 1) {
      double t[2];
      check(accp_calctune(L.device(), dpp, delta, t), "accp_calctune");
      std::cout << "Tunes: [" << fmt(t[0]) << "\t" << fmt(t[1]) << "], dpp=" << fmt(dpp) << std::endl;
    } else {
      double c[2];
      check(accp_calcchrom(L.device(), dpp, delta, c), "accp_calcchrom");
      std::cout << "Chromaticity: [" << fmt(c[0]) << "\t" << fmt(c[1]) << "], dpp=" << fmt(dpp) << std::endl;
    }
  }
}

// ---- Track block (process.cpp:1428-2220) -------------------------------------------
double find_num(const string &sel, const string &key, bool &found) {
  size_t p = sel.find(key);
  found = p != string::npos;
  if (!found) return 0;
  string s = sel.substr(p + key.size());
  s.erase(0, s.find('=') + 1);
  double v = 0;
  std::istringstream(s) >> v;
  return v;
}

// Box-Muller normals from a counter-based stream (Philox4x32-10 is the
// tracker's; here a splitmix64 counter suffices for initial beams).
struct Normals {
  uint64_t key, ctr = 0;
  explicit Normals(uint64_t seed) : key(seed) {}
  double uniform() {
    uint64_t z = key + 0x9E3779B97F4A7C15ull * ++ctr;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
  double next() {
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.141592653589793 * u2);
  }
};

// sigma and divergence of one plane of a Gaussian beam, given as emit_x/betx
// (betx from the ring-start Twiss when absent, 1 m when not positive or not
// finite) or as sigx/sigxp (process.cpp:1619-1735).  The reference reads betx
// at an offset taken from the wrong string (:1629-1633); its intended value
// is the number after "betx=", which is what is read here.
void gauss_plane(const string &sel, Line &L, int plane, double &sig, double &sigp) {
  const string emit = plane ? "emit_y" : "emit_x", bet = plane ? "bety" : "betx";
  bool f;
  const double e = find_num(sel, emit, f);
  if (!f) {
    sig = find_num(sel, plane ? "sigy" : "sigx", f);
    sigp = find_num(sel, plane ? "sigyp" : "sigxp", f);
    return;
  }
  double b = find_num(sel, bet, f);
  if (!f) {
    double tw[7];
    if (L.parse_only) b = 1.0;
    else {
      check(accp_calctwiss(L.device(), nullptr, 0, 0.0, 1e-8, tw), "accp_calctwiss");
      b = tw[3 + 2 * plane];
    }
  }
  if (!(b > 0) || !std::isfinite(b)) {
    b = 1.0;
    std::cout << "#" << bet << "=1.0 m is assumed" << std::endl;
  }
  sig = std::sqrt(b * e);
  sigp = std::sqrt(e / b);
}

std::vector<double> gen_distri(const string &line, uint64_t seed, Line &L) {
  std::cout << line << std::endl;
  string sel = trim(line.substr(line.find_first_of("=:") + 1));
  size_t pos = sel.find_first_of(" \t");
  if (pos == string::npos && !ieq(sel, "ZERO")) return std::vector<double>(6, 0.0);  // one particle at zero
  const string option = sel.substr(0, pos);
  std::vector<double> r;
  if (ieq(option, "ZERO")) {
    int np = 0;
    std::istringstream(sel.substr(sel.find('=') + 1)) >> np;
    r.assign(6 * (size_t)std::max(np, 0), 0.0);
  } else if (ieq(option, "LOAD")) {
    sel.erase(0, pos + 1);
    size_t q = sel.find_first_of(", \t\n");
    const string file = trim(sel.substr(0, q));
    std::cout << file << std::endl;
    string format = "PLAIN";
    string rest = q == string::npos ? "" : sel.substr(q + 1);
    size_t e = rest.find('=');
    if (e != string::npos && ieq(rest.substr(0, e), "FORMAT")) format = upper(trim(rest.substr(e + 1)));
    std::ifstream fin(file);
    if (!fin) throw std::runtime_error("cannot open distribution file " + file);
    string sl;
    if (format == "ACCEL-EXPORT")
      while (get_logical_line(fin, sl))
        if (sl.compare(0, 3, "$$$") == 0 && sl.find("Full export") != string::npos) {
          get_logical_line(fin, sl);
          break;
        }
    while (get_logical_line(fin, sl)) {
      if (sl.compare(0, 5, "*****") == 0) break;
      if (sl.size() <= 11) continue;
      std::istringstream ls(sl);
      double v[6];
      if (format == "ACCEL-EXPORT") {
        int turn;
        ls >> turn;
      }
      for (double &x : v) {
        string tok;
        ls >> tok;
        x = ieq(tok, "nan") || ieq(tok, "-nan") ? std::nan("") : std::stod(tok);
      }
      r.insert(r.end(), v, v + 6);
    }
  } else if (ieq(option, "COORDINATE")) {
    size_t a = sel.find('('), b = sel.find(')');
    string s = sel.substr(a + 1, b - a - 1);
    std::replace(s.begin(), s.end(), ',', ' ');
    std::istringstream ps(s);
    string part;
    while (std::getline(ps, part, ';')) {
      std::istringstream qs(part);
      double v[6] = {0, 0, 0, 0, 0, 0};
      for (double &x : v) qs >> x;
      r.insert(r.end(), v, v + 6);
    }
  } else if (ieq(option, "GAUSSIAN")) {
    sel.erase(0, pos + 1);
    bool f;
    int np = (int)find_num(sel, "NPar", f);
    if (!f) np = 1;
    double sg[6];
    gauss_plane(sel, L, 0, sg[0], sg[1]);
    gauss_plane(sel, L, 1, sg[2], sg[3]);
    sg[4] = find_num(sel, "sigz", f);
    sg[5] = find_num(sel, "sigdpp", f);
    Normals N(seed);
    r.resize(6 * (size_t)std::max(np, 0));
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < 6; ++j) r[6 * i + j] = sg[j] * N.next();
  } else if (ieq(option, "LINEAR")) {
    sel = trim(sel.substr(pos + 1));
    size_t e = sel.find('=');
    const string coord = trim(sel.substr(0, e));
    string s = trim(sel.substr(e + 1));
    std::replace(s.begin(), s.end(), ':', ' ');
    double st = 0, step = 0, ed = 0;
    std::istringstream(s) >> st >> step >> ed;
    if (step == 0) {
      std::cout << "interval cannot be zero" << std::endl;
      return r;
    }
    int np = 0;
    for (double v = st; (v <= ed && step > 0) || (v >= ed && step < 0); v += step) ++np;
    const char *names[6] = {"X", "XP", "Y", "YP", "Z", "DELTA"};
    int j = 0;
    for (int c = 0; c < 6; ++c)
      if (ieq(coord, names[c])) j = c;
    r.assign(6 * (size_t)np, 0.0);
    for (int i = 0; i < np; ++i) r[6 * i + j] = st + step * i;
  }
  return r;
}

struct Monitor {
  int interval = 1;
  double fraction = 1.0;
  int flag = 0;  // 1 Coordinate, 2 Stat1, 3 Stat2
};

void print_stat(const double *s, bool cross) {
  std::cout << (long long)s[0] << "\t" << (long long)s[1];
  for (int i = 0; i < 6; ++i) std::cout << "\t" << fmt_sci(s[2 + i]);
  for (int i = 0; i < 6; ++i) std::cout << "\t" << fmt_sci(s[8 + i]);
  if (cross)
    for (int i = 0; i < 6; ++i) std::cout << "\t" << fmt_sci(s[14 + i]);
  std::cout << std::endl;
}

void print_coords(const std::vector<double> &r, long long turn, double fraction) {
  const int np = (int)(r.size() / 6);
  int iskip = (int)(1.0 / fraction);
  iskip = std::max(1, std::min(iskip, np));
  for (int i = 0; i < np; i += iskip) {
    std::cout << turn;
    for (int j = 0; j < 6; ++j) std::cout << "\t" << fmt(r[6 * i + j]);
    std::cout << std::endl;
  }
}

void do_track(Line &L, std::vector<double> &r, const Monitor &mon, const string &line, uint64_t &seed) {
  string str_rad, str_rf, printfinal = "NO";
  int nturn = 0;
  for (const auto &kv : key_values(line)) {
    if (kv.first == "RADIATION") str_rad = kv.second;
    else if (kv.first == "RFCAVITY") str_rf = kv.second;
    else if (kv.first == "NTURN") nturn = std::stoi(kv.second);
    else if (kv.first == "PRINTFINAL") printfinal = kv.second;
    else if (kv.first == "SEED") seed = std::stoull(kv.second);
  }
  if (ieq(str_rad, "OFF")) L.set_rad(ACCP_RAD_OFF);
  else if (ieq(str_rad, "RAD")) L.set_rad(ACCP_RAD_DAMP);
  else if (ieq(str_rad, "RADQE")) L.set_rad(ACCP_RAD_QE);
  else if (ieq(str_rad, "LUMP")) L.set_rad(ACCP_RAD_LUMP);
  else if (ieq(str_rad, "LUMPDAMP")) L.set_rad(ACCP_RAD_LUMP_DAMP);
  else if (ieq(str_rad, "LUMPQE")) L.set_rad(ACCP_RAD_LUMP_QE);
  if (ieq(str_rf, "OFF")) L.set_on(ACCP_CAVITY, 0);
  else if (ieq(str_rf, "ON")) L.set_on(ACCP_CAVITY, 1);
  const int np = (int)(r.size() / 6);
  if (np <= 0 || nturn <= 0) return;
  const char *head[4] = {"", "$$$Track-Data: Coordinate**************************",
                         "$$$Track-Data: Stat1******************************",
                         "$$$Track-Data: Stat2******************************"};
  if (mon.flag) {
    std::cout << head[mon.flag] << std::endl;
    std::cout << "#" << str_rad << " " << str_rf << " " << nturn << std::endl;
    if (mon.flag == 1) {
      int iskip = std::max(1, std::min((int)(1.0 / mon.fraction), np));
      std::cout << "Np=" << np << ", Nprint=" << np / iskip << std::endl;
    } else {
      std::cout << "Np=" << np << std::endl;
    }
  }
  accp_lattice *lat = L.device();
  accp_track_opts o = track_opts(np);
  o.rng_seed = seed;
  if (mon.flag == 2 || mon.flag == 3) {
    // statistics at the start of every interval-th turn, on the device
    const int nrec = (nturn + mon.interval - 1) / mon.interval;
    std::vector<double> rec((size_t)nrec * ACCP_STAT_LEN);
    int32_t cnt = 0;
    o.mon_interval = mon.interval;
    o.mon_max = nrec;
    o.mon_out = rec.data();
    o.mon_count = &cnt;
    check(accp_track(lat, r.data(), np, nturn, 0, -1, &o), "accp_track");
    if (np >= 2)
      for (int k = 0; k < cnt; ++k) print_stat(&rec[(size_t)k * ACCP_STAT_LEN], mon.flag == 3);
  } else if (mon.flag == 1) {
    // coordinates at the start of every interval-th turn: track in pieces
    for (int t = 0; t < nturn; t += mon.interval) {
      print_coords(r, t, mon.fraction);
      o.turn0 = (uint64_t)t;
      check(accp_track(lat, r.data(), np, std::min(mon.interval, nturn - t), 0, -1, &o), "accp_track");
    }
  } else {
    check(accp_track(lat, r.data(), np, nturn, 0, -1, &o), "accp_track");
  }
  if (ieq(printfinal, "Full")) {
    std::cout << "*************************************" << std::endl;
    std::cout << "$$$Track-Data: Full export ************************" << std::endl;
    std::cout << "Np=" << np << ", Nprint=" << np << std::endl;
    for (int i = 0; i < np; ++i) {  // lossless: Distribution Load format=accel-export reads it back
      std::cout << nturn << "\t";
      for (int j = 0; j < 6; ++j) std::cout << lossless(r[6 * i + j]) << "\t";
      std::cout << std::endl;
    }
  } else if (ieq(printfinal, "Yes") || ieq(printfinal, "same")) {
    if (mon.flag == 1) print_coords(r, nturn, mon.fraction);
    else if ((mon.flag == 2 || mon.flag == 3) && np >= 2) {
      double s[ACCP_STAT_LEN];
      check(accp_beam_stat(r.data(), np, s), "accp_beam_stat");
      s[0] = nturn;
      print_stat(s, mon.flag == 3);
    }
  }
  std::cout << "*************************************" << std::endl;
}

// TRANSPORT: type=, famname=, index=n|[i j ...] (process.cpp:2108-2214):
// a copy of the beam passes elements [0, w0), then [w(k-1), wk) for each
// further watch point w, printing the particles at each.  The reference
// prints the untransported distribution (it writes r, having tracked its
// copy nr); the transported coordinates are printed here.
void cmd_transport(Line &L, const std::vector<double> &r, const string &line) {
  string s = line.substr(line.find_first_of("=:") + 1);
  std::replace(s.begin(), s.end(), ',', ' ');
  s = trim(s);
  string type, fam, indices;
  unsigned index = 0;
  for (int cnt = 0; cnt < 3 && !s.empty(); ++cnt) {
    size_t e = s.find('=');
    const string key = upper(trim(s.substr(0, e)));
    s = e == string::npos ? "" : trim(s.substr(e + 1));
    size_t b = s.find_first_of(" \t");
    if (key == "TYPE") type = s.substr(0, b);
    else if (key == "FAMNAME") fam = s.substr(0, b);
    else if (key == "INDEX") {
      if (!s.empty() && s.front() == '[') {
        b = s.find(']');
        indices = trim(s.substr(1, b - 1));
      } else {
        std::istringstream(s) >> index;
      }
    }
    s = b == string::npos ? "" : trim(s.substr(b + 1));
  }
  if (indices.empty()) std::cout << type << " " << fam << " " << index << std::endl;
  else std::cout << type << " " << fam << " " << indices << " " << std::endl;
  const std::vector<int32_t> list = elem_list(L, type, fam, index, indices);
  if (list.empty()) {
    std::cout << "No watch point found" << std::endl;
    return;
  }
  if (L.parse_only) return;
  const int np = (int)(r.size() / 6);
  std::vector<double> nr = r;
  std::cout << "$$$Track-Transport: ************************" << std::endl;
  int from = 0;
  for (int32_t w : list) {
    if (w < 0 || w > (int)L.e.size()) throw std::runtime_error("TRANSPORT: ring index out of range");
    if (w > from && np > 0) check(accp_track(L.device(), nr.data(), np, 1, from, w, nullptr), "accp_track");
    double spos = 0;
    for (int i = 0; i < w; ++i) spos += L.e[i].len;
    std::cout << "type=" << type << ", famname=" << fam << ", ring index=" << w << ", spos=" << fmt(spos)
              << std::endl;
    for (int i = 0; i < np; ++i) {
      std::cout << w << "\t";
      for (int j = 0; j < 6; ++j) std::cout << fmt(nr[6 * (size_t)i + j]) << "\t";
      std::cout << std::endl;
    }
    std::cout << "#------end of ring index=" << w << std::endl;
    from = w;
  }
  std::cout << "******************************" << std::endl;
}

void proc_track(std::istream &in, Line &L) {
  const std::vector<string> keys = {"DISTRIBUTION", "MONITOR", "DO_TRACK", "TRANSPORT", "END"};
  std::vector<double> r;
  Monitor mon;
  uint64_t seed = 0;
  string line;
  for (;;) {
    const int k = look_for(in, line, keys);
    if (k < 0 || keys[k] == "END") return;
    if (k == 0) {
      r = gen_distri(line, seed, L);
      const string up = upper(line);
      size_t p = up.find("SHIFT");
      if (p != string::npos && !r.empty()) {
        size_t a = up.find('(', p), b = up.find(')', a);
        string s = line.substr(a + 1, b - a - 1);
        std::replace(s.begin(), s.end(), ',', ' ');
        double d[6] = {0, 0, 0, 0, 0, 0};
        std::istringstream qs(s);
        for (double &x : d) qs >> x;
        for (size_t i = 0; i < r.size(); ++i) r[i] += d[i % 6];
      }
      if (up.find("PRINT") != string::npos) {
        std::cout << "$$$Init-Distribution*************" << std::endl;
        for (size_t i = 0; i < r.size(); i += 6) {
          for (int j = 0; j < 6; ++j) std::cout << fmt(r[i + j]) << " ";
          std::cout << std::endl;
        }
        std::cout << "***********************" << std::endl;
      }
    } else if (k == 1) {
      mon = Monitor();
      mon.flag = 1;
      for (const auto &kv : key_values(line)) {
        if (kv.first == "INTERVAL") mon.interval = std::max(1, std::stoi(kv.second));
        else if (kv.first == "FRACTION") mon.fraction = std::min(1.0, std::max(0.01, std::stod(kv.second)));
        else if (kv.first == "PRINTFLAG")
          mon.flag = ieq(kv.second, "STAT1") ? 2 : ieq(kv.second, "STAT2") ? 3 : 1;
      }
    } else if (k == 2) {
      if (L.parse_only) continue;
      do_track(L, r, mon, line, seed);
    } else if (k == 3) {
      cmd_transport(L, r, line);
    }
  }
}

// ---- DAMA block (process.cpp:2393-2500) --------------------------------------------
void proc_dama(std::istream &in, Line &L) {
  const std::vector<string> keys = {"TRACK_DA", "TRACK_MA", "APPLY_SETPT", "TRACK_DNUDJ", "END"};
  string line;
  for (;;) {
    const int k = look_for(in, line, keys);
    if (k < 0 || keys[k] == "END") return;
    if (k == 2) {
      cmd_apply_setpt(L, line);
      continue;
    }
    if (k == 3) {
      // TRACK_DNUDJ: xmax=, ymax=, nturn= (process.cpp:2587-2640): radiation
      // and cavities off, then the amplitude-dependent tune shift.  A key
      // left out takes ALine::trackdnudexy's default (accel.h:651) where the
      // reference reads an uninitialised variable.
      double xm = 0.004, ym = 0.003;
      int nturn = 512;
      for (const auto &kv : key_values(line)) {
        if (kv.first == "XMAX") xm = std::stod(kv.second);
        else if (kv.first == "YMAX") ym = std::stod(kv.second);
        else if (kv.first == "NTURN") nturn = std::stoi(kv.second);
      }
      std::cout << "#" << fmt(xm) << " " << fmt(ym) << " " << nturn << std::endl;
      if (L.parse_only) continue;
      L.set_rad(ACCP_RAD_OFF);
      L.set_on(ACCP_CAVITY, 0);
      double d[4], rows[36];
      check(accp_trackdnudexy(L.device(), nturn, xm, ym, d, rows), "accp_trackdnudexy");
      for (int i = 0; i < 12; ++i)  // the lines trackdnudexy prints (aline.cpp:744, :768)
        std::cout << fmt(rows[3 * i]) << "\t" << fmt(rows[3 * i + 1]) << "\t" << fmt(rows[3 * i + 2]) << std::endl;
      std::cout << "$$$dtuning coefficients" << std::endl;
      std::cout << fmt(d[0]) << " \t" << fmt(d[1]) << " \t" << fmt(d[2]) << " \t" << fmt(d[3]) << " \t" << std::endl;
      std::cout << "*************************" << std::endl;
      continue;
    }
    if (k == 1) {
      // TRACK_MA (process.cpp:2503-2580): locations = elements of type
      // [and famname], the first nmoni of them
      string type, fam;
      unsigned nmoni = 0;
      int nstep = 0, nturn = 0;
      double dppmin = 0, dppmax = 0;
      for (const auto &kv : key_values(line)) {
        if (kv.first == "TYPE") type = kv.second;
        else if (kv.first == "FAMNAME") fam = kv.second;
        else if (kv.first == "DPPMIN") dppmin = std::stod(kv.second);
        else if (kv.first == "DPPMAX") dppmax = std::stod(kv.second);
        else if (kv.first == "NMONI") nmoni = (unsigned)std::stoul(kv.second);
        else if (kv.first == "NSTEP") nstep = std::stoi(kv.second);
        else if (kv.first == "NTURN") nturn = std::stoi(kv.second);
      }
      std::cout << "#" << nmoni << " " << nstep << " " << nturn << " " << fmt(dppmin) << " " << fmt(dppmax)
                << std::endl;
      std::vector<int32_t> list = L.find(type, fam);
      if (nmoni > list.size()) nmoni = (unsigned)list.size();
      list.resize(nmoni);
      if (L.parse_only || list.empty()) continue;
      L.set_rad(ACCP_RAD_DAMP);
      L.set_on(ACCP_CAVITY, 1);
      std::vector<double> plus(nmoni), minus(nmoni);
      double avg = 0;
      check(accp_trackma(L.device(), dppmax, dppmin, nturn, nstep, list.data(), (int)nmoni, plus.data(),
                         minus.data(), &avg),
            "accp_trackma");
      std::cout << "$$$MA " << nturn << " turns, average MA=" << fmt(avg) << std::endl;
      for (unsigned i = 0; i < nmoni; ++i) std::cout << fmt(plus[i]) << " \t" << fmt(minus[i]) << " \t" << std::endl;
      std::cout << "******************************" << std::endl;
      continue;
    }
    int nray = 0, nstep = 0, nturn = 0;
    double xmin = 0, xmax = 0, ymax = 0;
    bool loss = false, neg = false;
    for (const auto &kv : key_values(line)) {
      if (kv.first == "XMIN") xmin = std::stod(kv.second);
      else if (kv.first == "XMAX") xmax = std::stod(kv.second);
      else if (kv.first == "YMAX") ymax = std::stod(kv.second);
      else if (kv.first == "NRAY") nray = std::stoi(kv.second);
      else if (kv.first == "NSTEP") nstep = std::stoi(kv.second);
      else if (kv.first == "NTURN") nturn = std::stoi(kv.second);
      else if (kv.first == "PRINT") loss = ieq(kv.second, "loss_turn");
      else if (kv.first == "WEIGHTMINUSSIDE") neg = ieq(kv.second, "true");
    }
    if (nray < 2 || nray > 51) std::cout << "Error: number of rays: " << nray << std::endl;
    std::cout << "#" << fmt(xmin) << " " << fmt(xmax) << " " << fmt(ymax) << " " << nray << " " << nstep << " "
              << nturn << " " << neg << std::endl;
    if (L.parse_only) continue;
    L.set_rad(ACCP_RAD_DAMP);
    L.set_on(ACCP_CAVITY, 1);
    const int nr = std::min(std::max(nray, 3), 51), ns = std::max(nstep, 2);
    double rx[51], ry[51], area = 0;
    std::vector<int32_t> lt((size_t)nr * ns);
    check(accp_trackda(L.device(), xmax, xmin, ymax, nturn, nray, nstep, neg, rx, ry, &area, lt.data()),
          "accp_trackda");
    if (loss) {
      // the reference's per-turn scan (aline.cpp:636-662) records the turn
      // after which a point is first seen lost, 0 for survivors; a point
      // lost in turn 0 is first recorded at turn 1
      std::cout << "$$$DA-Loss " << std::endl;
      const double dx = xmax - xmin;
      for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
          const double rad = xmin + dx * j / (ns - 1);
          const double xl = rad * std::cos(i * 3.1415926535897932384626 / (nr - 1));
          const double yl = rad * std::sin(i * 3.1415926535897932384626 / (nr - 1)) * ymax / xmax;
          const int t = lt[(size_t)i * ns + j];
          const int rep = t >= nturn ? 0 : (t == 0 ? (nturn > 1 ? 1 : 0) : t);
          std::cout << fmt(xl) << "\t" << fmt(yl) << "\t" << rep << std::endl;
        }
    }
    std::cout << "$$$DA-profile " << nturn << " turns, " << (neg ? "3/4 DA area=" : "DA area=") << fmt(area * 1.0e6)
              << " mm^2" << std::endl;
    for (int i = 0; i < nr; ++i) std::cout << fmt(rx[i]) << " \t" << fmt(ry[i]) << " \t" << std::endl;
    std::cout << "******************************" << std::endl;
  }
}

}  // namespace

int main(int argc, char **argv) {
  string cmdfile;
  bool parse_only = false;
  for (int i = 1; i < argc; ++i) {
    const string a = argv[i];
    if (a == "--precision" && i + 1 < argc) g_prec = std::max(1, std::min(17, std::atoi(argv[++i])));
    else if (a == "--specialized") g_specialized = true;
    else if (a == "--parse-only") parse_only = true;
    else cmdfile = a;
  }
  if (cmdfile.empty()) {
    std::cerr << "usage: accp-run [--precision N] [--specialized] [--parse-only] command_file" << std::endl;
    return 2;
  }
  std::ifstream in(cmdfile);
  if (!in) {
    std::cerr << "cannot open " << cmdfile << std::endl;
    return 2;
  }
  const std::vector<string> blocks = {"LATTICE", "OPTICS", "SURVEY", "DAMA", "TRACK"};
  Line L;
  L.parse_only = parse_only;
  try {
    string line;
    int k = look_for(in, line, blocks);
    if (k != 0) {
      std::cout << (k > 0 ? "The first block must be LATTICE." : "No LATTICE block is found.") << std::endl;
      return 1;
    }
    proc_lattice(in, L);
    std::cout << "Total elements: " << L.e.size() << std::endl;
    while ((k = look_for(in, line, blocks)) >= 0) {
      std::cout << blocks[k] << std::endl;
      switch (k) {
        case 0: proc_lattice(in, L); break;
        case 1: proc_optics(in, L); break;
        case 3: proc_dama(in, L); break;
        case 4: proc_track(in, L); break;
        default: {  // Survey: nothing on the tracking path
          std::vector<string> end = {"END"};
          look_for(in, line, end);
        }
      }
    }
  } catch (const std::exception &ex) {
    std::cout.flush();
    std::cerr << "accp-run: " << ex.what() << std::endl;
    return 1;
  }
  return 0;
}
