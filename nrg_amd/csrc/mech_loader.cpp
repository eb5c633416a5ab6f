// mech_loader.cpp -- mechanism, NASA-7 and molar-mass files read as NRG reads
// them (include/nrg_mech.h).  Host code only.
//
// Reading rules follow package_library/src/chemical_properties_class.f90
// (set_properties :117-181, reactions :289-451, sides :453-635, species
// :685-768) and package_library/src/thermophysical_properties_class.f90
// (:137-179, get_a_coeffs :1226-1311, molar masses :1313-1341).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nrg_mech.h"
#include "nrg_common.h"

using nrg::set_error;

struct nrg_mech_file {
  std::vector<std::string> species;
  std::vector<std::string> reaction_strings;
  std::vector<int32_t> reactions_type;
  std::vector<int32_t> chemical_coeffs;        // Fortran (4,R,2)
  std::vector<double> A, beta, E_act, A_low, beta_low, E_act_low;
  std::vector<double> Troe_coeffs;             // Fortran (R,4)
  std::vector<double> enhanced_efficiencies;   // Fortran (R,S)
  std::vector<double> a_coeffs;                // Fortran (7,2,S), x R
  std::vector<double> molar_masses;
};

namespace {

constexpr double R_GAS_J = 8.31446261815324;     // global_data.f90:36
constexpr double R_GAS_CAL = 1.98720425864083;   // :37

struct FormatError {
  std::string msg;
};

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\f' || c == '\v'; }

std::string strip(const std::string &s)
{
  size_t a = 0, b = s.size();
  while (a < b && is_space(s[a])) ++a;
  while (b > a && is_space(s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string rstrip(const std::string &s)
{
  size_t b = s.size();
  while (b > 0 && is_space(s[b - 1])) --b;
  return s.substr(0, b);
}

std::vector<std::string> read_lines(const char *path)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) throw FormatError{std::string("cannot open ") + path};
  std::vector<std::string> lines;
  std::string ln;
  while (std::getline(f, ln)) {
    while (!ln.empty() && (ln.back() == '\r' || ln.back() == '\n')) ln.pop_back();
    lines.push_back(ln);
  }
  return lines;
}

// A Fortran list-directed real: "4.8749+04" (exponent letter omitted) and D
// exponents are accepted; a trailing comma is dropped.
bool fortran_real(const std::string &tok, double &out)
{
  std::string t = strip(tok);
  while (!t.empty() && t.back() == ',') t.pop_back();
  // mantissa [+-]?(\d+\.?\d*|\.\d+) followed directly by [+-]\d+
  {
    size_t i = 0, n = t.size();
    if (i < n && (t[i] == '+' || t[i] == '-')) ++i;
    const size_t m0 = i;
    bool digits = false;
    while (i < n && std::isdigit((unsigned char)t[i])) ++i, digits = true;
    if (i < n && t[i] == '.') {
      ++i;
      while (i < n && std::isdigit((unsigned char)t[i])) ++i, digits = true;
    }
    if (digits && i > m0 && i < n && (t[i] == '+' || t[i] == '-') && i + 1 < n) {
      size_t j = i + 1;
      while (j < n && std::isdigit((unsigned char)t[j])) ++j;
      if (j == n) t = t.substr(0, i) + "E" + t.substr(i);
    }
  }
  for (char &c : t) {
    if (c == 'D') c = 'E';
    else if (c == 'd') c = 'e';
  }
  if (t.empty()) return false;
  // plain decimal forms only (no hex, no inf/nan spellings beyond the usual)
  for (char c : t)
    if (!(std::isdigit((unsigned char)c) || c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-'))
      return false;
  errno = 0;
  char *end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (end != t.c_str() + t.size()) return false;
  out = v;
  return true;
}

// One record split as list-directed input splits it: blank or comma
// separators, a quoted string is one item ('' inside quotes is a quote), an
// unquoted '/' ends the record.
std::vector<std::string> list_directed_items(const std::string &line)
{
  std::vector<std::string> items;
  size_t i = 0, n = line.size();
  while (i < n) {
    const char c = line[i];
    if (c == ' ' || c == '\t' || c == ',') {
      ++i;
      continue;
    }
    if (c == '/') break;
    if (c == '\'' || c == '"') {
      const char q = c;
      std::string buf;
      size_t j = i + 1;
      while (j < n) {
        if (line[j] == q) {
          if (j + 1 < n && line[j + 1] == q) {
            buf.push_back(q);
            j += 2;
            continue;
          }
          break;
        }
        buf.push_back(line[j]);
        ++j;
      }
      items.push_back(buf);
      i = j + 1;
      continue;
    }
    size_t j = i;
    while (j < n && line[j] != ' ' && line[j] != '\t' && line[j] != ',' && line[j] != '/') ++j;
    items.push_back(line.substr(i, j - i));
    i = j;
  }
  return items;
}

// Sequential records with read(*) / backspace.
struct Records {
  const std::vector<std::string> &lines;
  size_t pos = 0;
  explicit Records(const std::vector<std::string> &l) : lines(l) {}
  // list-directed read of n items; records without items are skipped;
  // false at end of file
  bool read_items(size_t n, std::vector<std::string> &got)
  {
    got.clear();
    while (got.size() < n) {
      if (pos >= lines.size()) return false;
      std::vector<std::string> it = list_directed_items(lines[pos]);
      ++pos;
      for (size_t k = 0; k < it.size() && got.size() < n; ++k) got.push_back(it[k]);
    }
    return true;
  }
  void backspace() { if (pos > 0) --pos; }
};

// read(unit,*,iostat=error) word, v(1:n): values convert in order, a bad item
// ends the read (remaining values unassigned); returns true on error.
bool read_values(Records &rec, size_t n, std::vector<double> &vals)
{
  vals.clear();
  std::vector<std::string> items;
  if (!rec.read_items(n + 1, items)) return true;
  for (size_t k = 1; k < items.size(); ++k) {
    double v;
    if (!fortran_real(items[k], v)) return true;
    vals.push_back(v);
  }
  return false;
}

std::vector<std::string> species_names(const std::vector<std::string> &lines)
{
  size_t k = 0;
  while (k < lines.size()) {
    std::vector<std::string> it = list_directed_items(lines[k]);
    if (!it.empty() && it[0] == "SPECIES") break;
    ++k;
  }
  ++k;
  std::vector<std::string> names;
  for (; k < lines.size(); ++k) {
    std::istringstream ss(lines[k]);
    std::string tok;
    while (ss >> tok) {
      if (tok == "END") return names;
      names.push_back(tok);
    }
  }
  throw FormatError{"SPECIES block has no END"};
}

// 1-based index in species + "M" (get_chemical_specie_index, :637-658); 0 if unknown
struct Index {
  const std::vector<std::string> &names;
  int operator()(const std::string &name) const
  {
    const std::string s = strip(name);
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == s) return (int)i + 1;
    if (s == "M") return (int)names.size() + 1;
    return 0;
  }
};

// One side of a reaction (:488-634): term count, up to three 1-based indices.
void parse_side(const std::string &side, const Index &index, int &count, int sp[3], bool &unknown)
{
  sp[0] = sp[1] = sp[2] = 0;
  unknown = false;
  size_t pos = side.find('+');
  if (pos == std::string::npos) {
    if (!side.empty() && side[0] == '2') {
      sp[0] = index(side.substr(1));
      sp[1] = sp[0];
      count = 2;
    } else {
      sp[0] = index(side);
      count = 1;
    }
    unknown = unknown || sp[0] == 0;
    return;
  }
  const std::string first = side.substr(0, pos);
  bool doubled = false;
  if (!first.empty() && first[0] == '2') {
    sp[0] = index(first.substr(1));
    sp[1] = sp[0];
    count = 2;
    doubled = true;
  } else {
    sp[0] = index(first);
    count = 1;
  }
  unknown = unknown || sp[0] == 0;
  const std::string rest = side.substr(pos + 1);
  pos = rest.find('+');
  if (pos == std::string::npos) {
    if (!rest.empty() && rest[0] == '2') {
      sp[1] = index(rest.substr(1));
      sp[2] = sp[1];
      count = 3;
      unknown = unknown || sp[1] == 0;
    } else if (doubled) {
      sp[2] = index(rest);
      count = 3;
      unknown = unknown || sp[2] == 0;
    } else {
      sp[1] = index(rest);
      count = 2;
      unknown = unknown || sp[1] == 0;
    }
  } else {
    sp[1] = index(rest.substr(0, pos));
    unknown = unknown || sp[1] == 0;
    sp[2] = index(rest.substr(pos + 1));
    unknown = unknown || sp[2] == 0;
    count = 3;
  }
}

// '=' -> 1, '=>' -> 2, '<=>' -> 3 (:327-374)
int split_reaction(const std::string &s, std::string &left, std::string &right)
{
  int t = 0;
  if (s.find('=') != std::string::npos) t = 1;
  if (s.find("=>") != std::string::npos) t = 2;
  if (s.find("<=") != std::string::npos) t = 3;
  if (t == 1) {
    const size_t p = s.find('=');
    left = s.substr(0, p);
    right = s.substr(p + 1);
  } else if (t == 2) {
    const size_t p = s.find("=>");
    left = s.substr(0, p);
    right = s.substr(p + 2);
  } else if (t == 3) {
    const size_t p = s.find("<=");
    left = s.substr(0, p);
    right = s.substr(p + 3);
  }
  return t;
}

void parse_mechanism(const char *path, double default_eff, const std::string &units, nrg_mech_file &m)
{
  const std::vector<std::string> lines = read_lines(path);
  m.species = species_names(lines);
  const int S = (int)m.species.size();
  const Index index{m.species};
  Records rec(lines);
  std::vector<std::string> items;
  for (;;) {
    if (!rec.read_items(1, items)) throw FormatError{"no REACTIONS section"};
    if (items[0] == "REACTIONS") break;
  }
  struct Rxn {
    int type;
    int left[4], right[4];
    double a, b, e, alo, blo, elo, troe[4];
    std::vector<double> eff;
    std::string str;
  };
  std::vector<Rxn> rx;
  bool enhanced = false;
  for (;;) {
    if (!rec.read_items(1, items)) break;
    std::string s = items[0];
    const std::string original = s;
    const size_t third_body = s.find("+M");
    const size_t fall_off = s.find("(+M)");
    if (fall_off != std::string::npos) {
      const std::string trimmed = rstrip(s);
      const size_t stop = trimmed.size() >= 4 ? trimmed.size() - 4 : 0;
      const std::string tail = (fall_off + 4 < stop) ? s.substr(fall_off + 4, stop - (fall_off + 4)) : "";
      s = s.substr(0, fall_off) + tail;
      enhanced = true;
    }
    std::string left, right;
    const int rtype = split_reaction(s, left, right);
    if (rtype == 0) continue;
    Rxn r{};
    bool unk_l, unk_r;
    int nl, nr;
    parse_side(strip(left), index, nl, r.left + 1, unk_l);
    parse_side(strip(right), index, nr, r.right + 1, unk_r);
    if (unk_l || unk_r) continue;
    r.left[0] = nl;
    r.right[0] = nr;
    if (rtype == 2) r.type = third_body != std::string::npos ? 6 : 5;
    else r.type = third_body != std::string::npos ? 1 : 0;
    if (third_body != std::string::npos) enhanced = true;
    r.eff.assign(S, default_eff);
    rec.backspace();
    // every read below is iostat= (:391-418); unassigned values stay 0 (:150-156)
    std::vector<double> v;
    bool error = read_values(rec, 3, v);
    r.a = v.size() > 0 ? v[0] : 0.0;
    r.b = v.size() > 1 ? v[1] : 0.0;
    r.e = v.size() > 2 ? v[2] : 0.0;
    bool eof = false;   // the file ends inside this reaction's records
    if (fall_off != std::string::npos) {
      if (!rec.read_items(1, items)) {
        eof = true;
      } else {
        error = false;
        if (items[0] == "LOW") {
          r.type = rtype == 2 ? 7 : 2;
          rec.backspace();
          error = read_values(rec, 3, v);
          r.alo = v.size() > 0 ? v[0] : 0.0;
          r.blo = v.size() > 1 ? v[1] : 0.0;
          r.elo = v.size() > 2 ? v[2] : 0.0;
        }
        if (!rec.read_items(1, items)) {
          eof = true;
        } else {
          error = false;
          if (items[0] == "TROE") {
            r.type = rtype == 2 ? 7 : 3;
            rec.backspace();
            error = read_values(rec, 4, v);
            for (size_t k = 0; k < 4; ++k) r.troe[k] = k < v.size() ? v[k] : 0.0;
          } else {
            rec.backspace();
          }
        }
      }
    }
    // efficiency record (:421-445), read only when the last read succeeded
    if (!error && !eof) {
      std::string tok;
      if (rec.read_items(1, items)) tok = items[0];
      if (tok.find('/') != std::string::npos && enhanced) {
        std::string rest = tok;
        while (!strip(rest).empty()) {
          const size_t p1 = rest.find('/');
          if (p1 == std::string::npos) break;
          const std::string sname = strip(rest.substr(0, p1));
          rest = rest.substr(p1 + 1);
          const size_t p2 = rest.find('/');
          const std::string value = p2 == std::string::npos ? rest : rest.substr(0, p2);
          const int sidx = index(sname);
          if (sidx > 0 && sidx <= S) {
            double val;
            if (!fortran_real(value, val)) throw FormatError{"bad efficiency in '" + tok + "'"};
            r.eff[sidx - 1] = val;
          }
          rest = p2 == std::string::npos ? std::string() : rest.substr(p2 + 1);
        }
      } else {
        rec.backspace();
      }
    }
    r.str = original;
    rx.push_back(std::move(r));
    if (eof) break;
  }
  double coef;
  if (units == "J.mol") coef = 1.0;
  else if (units == "kJ.mol") coef = 1000.0;
  else if (units == "cal.mol") coef = R_GAS_J / R_GAS_CAL;
  else if (units == "kcal.mol") coef = R_GAS_J / R_GAS_CAL * 1000.0;
  else throw FormatError{"unknown E_act_units '" + units + "'"};
  const int R = (int)rx.size();
  m.reactions_type.resize(R);
  m.chemical_coeffs.assign((size_t)2 * R * 4, 0);
  m.A.resize(R), m.beta.resize(R), m.E_act.resize(R);
  m.A_low.resize(R), m.beta_low.resize(R), m.E_act_low.resize(R);
  m.Troe_coeffs.assign((size_t)4 * R, 0.0);
  m.enhanced_efficiencies.assign((size_t)S * R, 0.0);
  m.reaction_strings.resize(R);
  for (int r = 0; r < R; ++r) {
    const Rxn &x = rx[r];
    m.reactions_type[r] = x.type;
    for (int c = 0; c < 4; ++c) {
      m.chemical_coeffs[(size_t)r * 4 + c] = x.left[c];
      m.chemical_coeffs[(size_t)R * 4 + (size_t)r * 4 + c] = x.right[c];
    }
    double A = x.a, A_low = x.alo;
    if (x.left[0] == 2) A = A * 1.0e-06;                 // :162-180
    if (x.left[0] == 3) A = A * 1.0e-12;
    if (x.left[0] + 1 == 2) A_low = A_low * 1.0e-06;
    if (x.left[0] + 1 == 3) A_low = A_low * 1.0e-12;
    m.A[r] = A;
    m.A_low[r] = A_low;
    m.beta[r] = x.b;
    m.beta_low[r] = x.blo;
    m.E_act[r] = x.e * coef;
    m.E_act_low[r] = x.elo * coef;
    for (int k = 0; k < 4; ++k) m.Troe_coeffs[(size_t)k * R + r] = x.troe[k];
    for (int s = 0; s < S; ++s) m.enhanced_efficiencies[(size_t)s * R + r] = x.eff[s];
    m.reaction_strings[r] = x.str;
  }
}

// Tokenizer of get_a_coeffs (:1259-1304): a number ends at a character that
// cannot be part of one, a '-' not following 'E' starts a new one, and the last
// character of the trimmed record ends one.
std::vector<std::string> nasa7_tokens(const std::string &line)
{
  const std::string s = rstrip(line);
  static const std::string classes = "-+E0123456789.";
  std::vector<std::string> out;
  const int n = (int)s.size();
  int prev = -1, beginning = 0;
  for (int i = 0; i < n; ++i) {
    const size_t f = classes.find(s[i]);
    const int indx = f == std::string::npos ? 0 : (int)f + 1;
    const bool boundary = (indx == 0 && prev > 0) || (indx == 1 && prev != 3 && prev > 0) ||
                          (indx != 0 && i == n - 1);
    if (boundary) {
      if (i != n - 1) out.push_back(s.substr(beginning, i - beginning));
      else out.push_back(s.substr(beginning, i + 1 - beginning));
      beginning = i;
    }
    prev = indx;
  }
  return out;
}

void a_coeffs_for(const std::string &name, const std::vector<std::string> &lines, double a[2][7])
{
  for (int st = 0; st < 2; ++st)
    for (int i = 0; i < 7; ++i) a[st][i] = 0.0;
  size_t k = 0;
  while (k < lines.size()) {
    const std::vector<std::string> it = list_directed_items(lines[k]);
    ++k;
    if (it.empty()) continue;
    if (it[0] == "END") break;
    if (it[0] != name) continue;
    int num = 0, i1 = 1, i2 = 2;
    while (num < 16) {
      if (k >= lines.size()) throw FormatError{"unsuccessful specific heat capacity reading " + name};
      const std::string &line = lines[k];
      ++k;
      for (const std::string &text : nasa7_tokens(line)) {
        if (num == 5 || num == 11) {
          ++num;
          continue;
        }
        if (num == 16) break;
        ++num;
        double v;
        if (!fortran_real(text, v)) throw FormatError{"bad NASA-7 value '" + text + "' for " + name};
        a[i2 - 1][i1 - 1] = v;
        ++i1;
        if (num == 8) i1 = 1, i2 = 1;
      }
    }
    // a later block of the same name overwrites this one, as in the reference
  }
}

void parse_thermo(const char *thermo_path, const char *mm_path, nrg_mech_file &m)
{
  const std::vector<std::string> tl = read_lines(thermo_path);
  const size_t S = m.species.size();
  m.a_coeffs.assign(S * 14, 0.0);
  for (size_t s = 0; s < S; ++s) {
    double a[2][7];
    a_coeffs_for(m.species[s], tl, a);
    for (int st = 0; st < 2; ++st)
      for (int i = 0; i < 7; ++i) m.a_coeffs[s * 14 + st * 7 + i] = a[st][i] * R_GAS_J;
  }
  const std::vector<std::string> ml = read_lines(mm_path);
  m.molar_masses.assign(S, 0.0);
  for (size_t s = 0; s < S; ++s) {
    bool found = false;
    for (const std::string &line : ml) {
      const std::vector<std::string> it = list_directed_items(line);
      if (it.size() < 2) throw FormatError{"molar mass is not defined for specie: " + m.species[s]};
      if (it[0] == m.species[s]) {
        if (!fortran_real(it[1], m.molar_masses[s]))
          throw FormatError{"bad molar mass for " + m.species[s]};
        found = true;
        break;
      }
    }
    if (!found) throw FormatError{"molar mass is not defined for specie: " + m.species[s]};
  }
}

// ---- task set-up: the namelist files that name the data files -------------
//
// NRG's main program builds the chemistry and the thermochemistry from two
// namelist files of its task set-up folder (computing_module/src/current_build/
// main.f90:170-171; file names package_library/src/global_data.f90:12-18):
//   chemical_data.inf         &chemical_properties chemical_mechanism_file_name,
//                             default_enhanced_efficiencies, E_act_units /
//                             (chemical_properties_class.f90:77-97)
//   thermophysical_setup.inf  &thermo_properties thermo_data_file_name,
//                             transport_data_file_name, molar_masses_data_file_name /
//                             (thermophysical_properties_class.f90:97-134)
// and opens the named files under chemical_mechanisms/ and thermophysical_data/
// of the same folder (chemical_properties_class.f90:135,
// thermophysical_properties_class.f90:161-163).

// One namelist group read from a file: lower-cased names -> value text
// (character values unquoted, Fortran's doubled quotes undone, trailing blanks
// dropped, as a character variable compares and as OPEN uses it).
std::vector<std::pair<std::string, std::string>> read_namelist(const char *path, const std::string &group)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) throw FormatError{std::string("cannot open ") + path};
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string txt = ss.str();
  auto lower = [](std::string v) {
    for (char &c : v) c = (char)std::tolower((unsigned char)c);
    return v;
  };
  // find "&group" (case-insensitive) outside comments
  size_t i = 0;
  const std::string want = lower(group);
  bool found = false;
  while (i < txt.size()) {
    if (txt[i] == '!') {
      while (i < txt.size() && txt[i] != '\n') ++i;
      continue;
    }
    if (txt[i] == '&') {
      size_t j = i + 1;
      while (j < txt.size() && (std::isalnum((unsigned char)txt[j]) || txt[j] == '_')) ++j;
      if (lower(txt.substr(i + 1, j - i - 1)) == want) {
        i = j;
        found = true;
        break;
      }
    }
    ++i;
  }
  if (!found) throw FormatError{std::string("no &") + group + " namelist group in " + path};
  std::vector<std::pair<std::string, std::string>> items;
  auto skip_sep = [&]() {
    while (i < txt.size()) {
      const char c = txt[i];
      if (is_space(c) || c == ',') {
        ++i;
      } else if (c == '!') {
        while (i < txt.size() && txt[i] != '\n') ++i;
      } else {
        break;
      }
    }
  };
  for (;;) {
    skip_sep();
    if (i >= txt.size()) throw FormatError{std::string("namelist &") + group + " not terminated in " + path};
    if (txt[i] == '/') break;
    if (txt[i] == '&') {                      // "&end"
      size_t j = i + 1;
      while (j < txt.size() && std::isalpha((unsigned char)txt[j])) ++j;
      if (lower(txt.substr(i + 1, j - i - 1)) == "end") break;
      throw FormatError{std::string("unexpected '&' in namelist &") + group + " of " + path};
    }
    size_t j = i;
    while (j < txt.size() && (std::isalnum((unsigned char)txt[j]) || txt[j] == '_')) ++j;
    if (j == i) throw FormatError{std::string("bad item in namelist &") + group + " of " + path};
    const std::string name = lower(txt.substr(i, j - i));
    i = j;
    while (i < txt.size() && is_space(txt[i])) ++i;
    if (i >= txt.size() || txt[i] != '=')
      throw FormatError{"expected '=' after " + name + " in namelist &" + group + " of " + path};
    ++i;
    while (i < txt.size() && is_space(txt[i])) ++i;
    std::string val;
    if (i < txt.size() && (txt[i] == '\'' || txt[i] == '"')) {
      const char q = txt[i++];
      for (;;) {
        if (i >= txt.size()) throw FormatError{"unterminated string for " + name + " in " + path};
        if (txt[i] == q) {
          if (i + 1 < txt.size() && txt[i + 1] == q) {
            val += q;
            i += 2;
            continue;
          }
          ++i;
          break;
        }
        val += txt[i++];
      }
      val = rstrip(val);
    } else {
      while (i < txt.size() && !is_space(txt[i]) && txt[i] != ',' && txt[i] != '/' && txt[i] != '!')
        val += txt[i++];
    }
    items.emplace_back(name, val);
  }
  return items;
}

std::string namelist_value(const std::vector<std::pair<std::string, std::string>> &items,
                           const std::string &name, const char *path)
{
  std::string v;
  bool have = false;
  for (const auto &kv : items)
    if (kv.first == name) {            // a later assignment of the same name wins
      v = kv.second;
      have = true;
    }
  if (!have) throw FormatError{name + " is not set in " + path};
  return v;
}

}  // namespace

extern "C" {

int nrg_mech_load_task_setup(const char *task_setup_dir, nrg_mech_file **out)
{
  if (!out) return set_error(NRG_CHEM_E_ARG, "out is NULL");
  *out = nullptr;
  if (!task_setup_dir) return set_error(NRG_CHEM_E_ARG, "task set-up folder is NULL");
  const std::string dir(task_setup_dir);
  std::string mech, thermo, molar, units;
  double eff = 0.0;
  try {
    const std::string chem_inf = dir + "/chemical_data.inf";
    const std::string thermo_inf = dir + "/thermophysical_setup.inf";
    const auto c = read_namelist(chem_inf.c_str(), "chemical_properties");
    mech = namelist_value(c, "chemical_mechanism_file_name", chem_inf.c_str());
    units = namelist_value(c, "e_act_units", chem_inf.c_str());
    const std::string e = namelist_value(c, "default_enhanced_efficiencies", chem_inf.c_str());
    if (!fortran_real(e, eff))
      throw FormatError{"default_enhanced_efficiencies '" + e + "' is not a real in " + chem_inf};
    const auto t = read_namelist(thermo_inf.c_str(), "thermo_properties");
    thermo = namelist_value(t, "thermo_data_file_name", thermo_inf.c_str());
    molar = namelist_value(t, "molar_masses_data_file_name", thermo_inf.c_str());
  } catch (const FormatError &e) {
    return set_error(NRG_CHEM_E_FORMAT, "%s", e.msg.c_str());
  } catch (const std::exception &e) {
    return set_error(NRG_CHEM_E_FORMAT, "%s", e.what());
  }
  const std::string mp = dir + "/chemical_mechanisms/" + mech;
  const std::string tp = dir + "/thermophysical_data/" + thermo;
  const std::string wp = dir + "/thermophysical_data/" + molar;
  return nrg_mech_load(mp.c_str(), tp.c_str(), wp.c_str(), eff, units.c_str(), out);
}

int nrg_mech_load(const char *mechanism_path, const char *thermo_path, const char *molar_masses_path,
                  double default_enhanced_efficiencies, const char *E_act_units, nrg_mech_file **out)
{
  if (!out) return set_error(NRG_CHEM_E_ARG, "out is NULL");
  *out = nullptr;
  if (!mechanism_path || !thermo_path || !molar_masses_path)
    return set_error(NRG_CHEM_E_ARG, "a file path is NULL");
  nrg_mech_file *m = new (std::nothrow) nrg_mech_file();
  if (!m) return set_error(NRG_CHEM_E_ARG, "out of host memory");
  try {
    parse_mechanism(mechanism_path, default_enhanced_efficiencies,
                    E_act_units ? std::string(E_act_units) : std::string("cal.mol"), *m);
    parse_thermo(thermo_path, molar_masses_path, *m);
  } catch (const FormatError &e) {
    delete m;
    return set_error(NRG_CHEM_E_FORMAT, "%s", e.msg.c_str());
  } catch (const std::exception &e) {
    delete m;
    return set_error(NRG_CHEM_E_FORMAT, "%s", e.what());
  }
  *out = m;
  return NRG_CHEM_OK;
}

int nrg_mech_view(const nrg_mech_file *m, nrg_chem_mech *v)
{
  if (!m || !v) return set_error(NRG_CHEM_E_ARG, "mechanism or view is NULL");
  v->species_number = (int32_t)m->species.size();
  v->reactions_number = (int32_t)m->reactions_type.size();
  v->reactions_type = m->reactions_type.data();
  v->chemical_coeffs = m->chemical_coeffs.data();
  v->A = m->A.data();
  v->beta = m->beta.data();
  v->E_act = m->E_act.data();
  v->A_low = m->A_low.data();
  v->beta_low = m->beta_low.data();
  v->E_act_low = m->E_act_low.data();
  v->Troe_coeffs = m->Troe_coeffs.data();
  v->enhanced_efficiencies = m->enhanced_efficiencies.data();
  v->a_coeffs = m->a_coeffs.data();
  v->molar_masses = m->molar_masses.data();
  return NRG_CHEM_OK;
}

const char *nrg_mech_species_name(const nrg_mech_file *m, int32_t s)
{
  if (!m || s < 0 || (size_t)s >= m->species.size()) return nullptr;
  return m->species[s].c_str();
}

const char *nrg_mech_reaction_string(const nrg_mech_file *m, int32_t r)
{
  if (!m || r < 0 || (size_t)r >= m->reaction_strings.size()) return nullptr;
  return m->reaction_strings[r].c_str();
}

void nrg_mech_free(nrg_mech_file *m) { delete m; }

}  // extern "C"
