// Structure readers (xyz in Angstrom, Turbomole coord in bohr) and a JSON
// writer for singlepoint results: the I/O around the singlepoint entry
// point.  The reference has none (its main builds molecules in code,
// src/main.cu); the formats are the ones tblite reads through mctc-lib and the
// JSON object mirrors tblite's result fields.  Host code only.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "tblite_amd.h"

namespace {

constexpr double kAngstromToBohr = 1.0 / 0.52917721090380;

const char* const kSymbols[] = {
    "",   "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si",
    "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
    "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru",
    "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr",
    "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",
    "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn"};
constexpr int kMaxZ = 86;

// Element symbol (any case) or atomic number -> Z; 0 if unknown.
int parse_element(const std::string& tok)
{
  if (tok.empty()) return 0;
  if (std::isdigit((unsigned char)tok[0])) {
    const int z = std::atoi(tok.c_str());
    return (z >= 1 && z <= kMaxZ) ? z : 0;
  }
  std::string s;
  for (char c : tok)
    if (std::isalpha((unsigned char)c)) s += c;
  if (s.empty() || s.size() > 2) return 0;
  s[0] = (char)std::toupper((unsigned char)s[0]);
  if (s.size() == 2) s[1] = (char)std::tolower((unsigned char)s[1]);
  for (int z = 1; z <= kMaxZ; ++z)
    if (s == kSymbols[z]) return z;
  return 0;
}

struct Mol {
  std::vector<int> num;
  std::vector<double> xyz;  // bohr
  double charge = 0.0;
  int uhf = 0;
};

bool read_xyz(std::istream& in, Mol& m)
{
  std::string line;
  if (!std::getline(in, line)) return false;
  const int nat = std::atoi(line.c_str());
  if (nat < 1) return false;
  if (!std::getline(in, line)) return false;  // comment
  for (int i = 0; i < nat; ++i) {
    if (!std::getline(in, line)) return false;
    std::istringstream ls(line);
    std::string el;
    double x, y, z;
    if (!(ls >> el >> x >> y >> z)) return false;
    const int zz = parse_element(el);
    if (zz == 0) return false;
    m.num.push_back(zz);
    m.xyz.push_back(x * kAngstromToBohr);
    m.xyz.push_back(y * kAngstromToBohr);
    m.xyz.push_back(z * kAngstromToBohr);
  }
  return true;
}

// $coord block: "x y z element" per line in bohr, until the next "$" group;
// "$coord angs" / "$coord frac" variants are not accepted.  A "$eht" group
// may carry charge=<q> and unpaired=<n>.
bool read_coord(std::istream& in, Mol& m)
{
  std::string line;
  bool in_coord = false, seen = false;
  while (std::getline(in, line)) {
    const size_t p = line.find_first_not_of(" \t");
    if (p == std::string::npos) continue;
    if (line[p] == '$') {
      in_coord = false;
      std::istringstream ls(line.substr(p));
      std::string key;
      ls >> key;
      if (key == "$coord") {
        std::string unit;
        if (ls >> unit && unit != "bohr") return false;
        in_coord = true;
        seen = true;
      } else if (key == "$eht") {
        std::string tok;
        while (ls >> tok) {
          if (tok.rfind("charge=", 0) == 0) m.charge = std::atof(tok.c_str() + 7);
          if (tok.rfind("unpaired=", 0) == 0) m.uhf = std::atoi(tok.c_str() + 9);
        }
      } else if (key == "$end") {
        break;
      }
      continue;
    }
    if (!in_coord) continue;
    std::istringstream ls(line);
    double x, y, z;
    std::string el;
    if (!(ls >> x >> y >> z >> el)) return false;
    const int zz = parse_element(el);
    if (zz == 0) return false;
    m.num.push_back(zz);
    m.xyz.push_back(x);
    m.xyz.push_back(y);
    m.xyz.push_back(z);
  }
  return seen && !m.num.empty();
}

bool ends_with(const std::string& s, const std::string& suffix)
{
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

void write_array(std::ostream& o, const char* key, const double* a, int n, bool last = false)
{
  o << "  \"" << key << "\": [";
  for (int i = 0; i < n; ++i) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.16e", a[i]);
    o << (i ? ", " : "") << buf;
  }
  o << "]" << (last ? "\n" : ",\n");
}

}  // namespace

extern "C" {

int tblite_amd_read_structure_file(const char* path, const char* format, int capacity, int* nat,
                                   int* num, double* xyz, double* charge, int* uhf)
{
  if (!path || !nat || capacity < 0) return TBLITE_AMD_ERR_ARGUMENT;
  std::ifstream in(path);
  if (!in) return TBLITE_AMD_ERR_ARGUMENT;
  std::string fmt = format ? format : "";
  const std::string p(path);
  if (fmt.empty()) fmt = ends_with(p, ".xyz") ? "xyz" : "coord";
  Mol m;
  bool ok = false;
  if (fmt == "xyz") ok = read_xyz(in, m);
  else if (fmt == "coord") ok = read_coord(in, m);
  if (!ok) return TBLITE_AMD_ERR_ARGUMENT;
  *nat = (int)m.num.size();
  if (charge) *charge = m.charge;
  if (uhf) *uhf = m.uhf;
  if (*nat > capacity) return TBLITE_AMD_OK;  // count only: call again with room
  if (num) std::memcpy(num, m.num.data(), sizeof(int) * m.num.size());
  if (xyz) std::memcpy(xyz, m.xyz.data(), sizeof(double) * m.xyz.size());
  return TBLITE_AMD_OK;
}

int tblite_amd_write_json(const char* path, int nat, const int* num, const double* xyz,
                          double energy, int iterations, const double* qat, const double* dpat,
                          const double* qpat, int nao, const double* emo, const double* focc)
{
  if (!path || nat < 1 || !num || !xyz || !std::isfinite(energy) || nao < 0)
    return TBLITE_AMD_ERR_ARGUMENT;
  std::ofstream o(path);
  if (!o) return TBLITE_AMD_ERR_ARGUMENT;
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.16e", energy);
  o << "{\n  \"version\": \"tblite_amd 1\",\n  \"method\": \"GFN2-xTB\",\n";
  o << "  \"energy\": " << buf << ",\n";
  o << "  \"scf-iterations\": " << iterations << ",\n";
  o << "  \"numbers\": [";
  for (int i = 0; i < nat; ++i) o << (i ? ", " : "") << num[i];
  o << "],\n";
  write_array(o, "positions", xyz, 3 * nat);
  if (qat) write_array(o, "charges", qat, nat);
  if (dpat) write_array(o, "atomic-dipoles", dpat, 3 * nat);
  if (qpat) write_array(o, "atomic-quadrupoles", qpat, 6 * nat);
  if (emo) write_array(o, "orbital-energies", emo, nao);
  if (focc) write_array(o, "orbital-occupations", focc, nao);
  // molecular dipole moment sum_a (q_a r_a + mu_a), atomic units
  double dip[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < nat && qat; ++a)
    for (int k = 0; k < 3; ++k) dip[k] += qat[a] * xyz[3 * a + k] + (dpat ? dpat[3 * a + k] : 0.0);
  write_array(o, "dipole", dip, 3, true);
  o << "}\n";
  return o ? TBLITE_AMD_OK : TBLITE_AMD_ERR_ARGUMENT;
}

}  // extern "C"
