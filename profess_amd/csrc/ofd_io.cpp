// On-disk formats of PROFESS inputs and densities, read and written on the
// host (no device work):
//   .ion geometry files   ReadGeometry  (ReadIonFile.f90:80-705)
//   .recpot tables        ReadPseudo's text reading (ReadIonFile.f90:708-930)
//   density files         Print3DArray (Output.f90:617-873, unscaled form)
//                         ReadDensity  (ReadIonFile.f90:1054-1314)
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ofd.h"

namespace {

constexpr double kBohr = 0.529177208607388;   // Angstrom per bohr (Constants.f90)
constexpr double kPi = 3.141592653589793238462643;

void set_err(char* err, int errlen, const std::string& msg) {
  if (err && errlen > 0) {
    std::snprintf(err, (size_t)errlen, "%s", msg.c_str());
  }
}

std::string upper(std::string s) {
  for (char& ch : s) ch = (char)std::toupper((unsigned char)ch);
  return s;
}

// whitespace-separated fields of a line (list-directed reading)
std::vector<std::string> fields(const std::string& line) {
  std::vector<std::string> out;
  std::istringstream is(line);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

bool to_double(const std::string& s, double* v) {
  // Fortran list-directed input also accepts a D exponent
  std::string t = s;
  for (char& ch : t)
    if (ch == 'd' || ch == 'D') ch = 'E';
  char* end = nullptr;
  *v = std::strtod(t.c_str(), &end);
  return end && *end == '\0' && end != t.c_str();
}

// 3x3 inverse of a (row i, column j)
void inverse3(const double a[3][3], double inv[3][3]) {
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                     a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
}

}  // namespace

static int read_geometry(const char* path, ofd_geometry* g, char* err, int errlen) {
  if (!path || !g) {
    set_err(err, errlen, "null argument");
    return -1;
  }
  std::memset(g, 0, sizeof(*g));
  std::ifstream in(path);
  if (!in) {
    set_err(err, errlen, std::string("cannot open ") + path);
    return -1;
  }
  std::vector<std::string> lines;
  for (std::string ln; std::getline(in, ln);) lines.push_back(ln);
  double cell[3][3] = {{0}};   // cell[i][j]: component i of lattice vector j (cellReal), Angstrom
  bool found_lattice = false, found_ions = false, cart = false, in_bohr = false;
  std::vector<std::string> names;          // species in order of first appearance
  std::vector<int> elem;                   // per ion, species index
  std::vector<double> coord;               // per ion, 3 values as read
  std::vector<std::string> pseudo;
  size_t i = 0;
  while (i < lines.size()) {
    auto f = fields(lines[i++]);
    if (f.empty() || f[0][0] == '#') continue;
    if (upper(f[0]) != "%BLOCK" || f.size() < 2) continue;
    const std::string cmd = upper(f[1]);
    if (cmd == "LATTICE_CART") {
      // READ(unit,*) cell%cellReal: nine values, column-major (a vector per row)
      std::vector<double> v;
      while (v.size() < 9 && i < lines.size()) {
        for (auto& t : fields(lines[i++])) {
          double x;
          if (!to_double(t, &x)) {
            set_err(err, errlen, "LATTICE_CART: not a number: " + t);
            return -2;
          }
          v.push_back(x);
        }
      }
      if (v.size() < 9) {
        set_err(err, errlen, "LATTICE_CART: fewer than nine values");
        return -2;
      }
      for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 3; ++c) cell[c][j] = v[3 * j + c];
      found_lattice = true;
    } else if (cmd == "LATTICE_ABC") {
      std::vector<double> v;
      while (v.size() < 6 && i < lines.size())
        for (auto& t : fields(lines[i++])) {
          double x;
          if (to_double(t, &x)) v.push_back(x);
        }
      if (v.size() < 6) {
        set_err(err, errlen, "LATTICE_ABC: fewer than six values");
        return -2;
      }
      const double a = v[0], b = v[1], c = v[2];
      const double al = v[3] * kPi / 180.0, be = v[4] * kPi / 180.0, ga = v[5] * kPi / 180.0;
      std::memset(cell, 0, sizeof(cell));
      cell[0][0] = a;
      cell[0][1] = b * std::cos(ga);
      cell[1][1] = b * std::sin(ga);
      cell[0][2] = c * std::cos(be);
      cell[1][2] = c * (std::cos(al) - std::cos(be) * std::cos(ga)) / std::sin(ga);
      const double c2 = c * c - cell[0][2] * cell[0][2] - cell[1][2] * cell[1][2];
      if (!(c2 > 0.0) || !(std::sin(ga) != 0.0)) {
        set_err(err, errlen, "LATTICE_ABC: lengths and angles do not form a cell");
        return -2;
      }
      cell[2][2] = std::sqrt(c2);
      found_lattice = true;
    } else if (cmd == "POSITIONS_FRAC" || cmd == "POSITIONS_CART" || cmd == "POSITIONS_ABS" ||
               cmd == "POSITIONS_BOHR") {
      if (found_ions) break;
      found_ions = true;
      cart = cmd == "POSITIONS_CART" || cmd == "POSITIONS_ABS";
      in_bohr = cmd == "POSITIONS_BOHR";
      while (i < lines.size()) {
        auto p = fields(lines[i++]);
        if (p.empty() || p[0][0] == '#') continue;
        if (upper(p[0]).compare(0, 4, "%END") == 0) break;
        double x[3];
        if (p.size() < 4 || !to_double(p[1], &x[0]) || !to_double(p[2], &x[1]) || !to_double(p[3], &x[2])) break;
        int e = -1;
        for (size_t k = 0; k < names.size(); ++k)
          if (names[k] == p[0]) e = (int)k;
        if (e < 0) {
          names.push_back(p[0]);
          e = (int)names.size() - 1;
        }
        elem.push_back(e);
        coord.insert(coord.end(), x, x + 3);
      }
      pseudo.assign(names.size(), std::string());
    } else if (cmd == "SPECIES_POT") {
      while (i < lines.size()) {
        auto p = fields(lines[i++]);
        if (p.empty() || p[0][0] == '#') continue;
        if (upper(p[0]).compare(0, 4, "%END") == 0) break;
        for (size_t k = 0; k < names.size(); ++k)
          if (names[k] == p[0] && p.size() > 1) pseudo[k] = p[1];
      }
    } else {
      // other blocks (SPECIES_CORE, SPECIES_RHOA, ION_OPTIMISATION, ...) are
      // skipped up to their %END
      while (i < lines.size()) {
        auto p = fields(lines[i++]);
        if (!p.empty() && upper(p[0]).compare(0, 4, "%END") == 0) break;
      }
    }
  }
  if (!found_lattice) {
    set_err(err, errlen, "no LATTICE_CART or LATTICE_ABC block");
    return -2;
  }
  if (!found_ions || elem.empty()) {
    set_err(err, errlen, "no POSITIONS block");
    return -2;
  }
  // handedness (ReadIonFile.f90:640-648): (a1 x a2) . a3 must be positive
  const double cx = cell[1][0] * cell[2][1] - cell[2][0] * cell[1][1];
  const double cy = cell[2][0] * cell[0][1] - cell[0][0] * cell[2][1];
  const double cz = cell[0][0] * cell[1][1] - cell[1][0] * cell[0][1];
  const double handed = cx * cell[0][2] + cy * cell[1][2] + cz * cell[2][2];
  if (handed < 0.0) {
    set_err(err, errlen, "left-handed lattice vectors");
    return -2;
  }
  if (!(handed > 0.0)) {
    set_err(err, errlen, "degenerate lattice vectors (zero volume)");
    return -2;
  }
  const int nion = (int)elem.size();
  // Cartesian (Angstrom) or bohr positions to fractional (ReadIonFile.f90:515-530)
  if (cart || in_bohr) {
    double inv[3][3];
    inverse3(cell, inv);
    for (int a = 0; a < nion; ++a) {
      double x[3] = {coord[3 * a], coord[3 * a + 1], coord[3 * a + 2]};
      if (in_bohr)
        for (double& v : x) v *= kBohr;
      for (int r = 0; r < 3; ++r) coord[3 * a + r] = inv[r][0] * x[0] + inv[r][1] * x[1] + inv[r][2] * x[2];
    }
  }
  for (double& v : coord) v = v - std::floor(v);   // coord - FLOOR(coord)
  // ions grouped by species in order of first appearance, stable within a
  // species (the exchange sort of ReadIonFile.f90:556-577 gives this order)
  std::vector<int> order(nion);
  for (int a = 0; a < nion; ++a) order[a] = a;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return elem[x] < elem[y]; });
  const int ns = (int)names.size();
  g->nion = nion;
  g->nspecies = ns;
  for (int j = 0; j < 3; ++j)
    for (int c = 0; c < 3; ++c) g->cell[3 * j + c] = cell[c][j] / kBohr;
  g->type_start = (int*)std::calloc((size_t)ns + 1, sizeof(int));
  g->frac = (double*)std::calloc(3 * (size_t)nion, sizeof(double));
  g->file_index = (int*)std::calloc((size_t)nion, sizeof(int));
  g->species = (ofd_name*)std::calloc((size_t)ns, sizeof(ofd_name));
  g->pseudo_file = (ofd_path*)std::calloc((size_t)ns, sizeof(ofd_path));
  if (!g->type_start || !g->frac || !g->file_index || !g->species || !g->pseudo_file) {
    ofd_free_geometry(g);
    set_err(err, errlen, "out of memory");
    return -1;
  }
  for (int a = 0; a < nion; ++a) {
    const int o = order[a];
    g->file_index[a] = o;
    for (int c = 0; c < 3; ++c) g->frac[3 * a + c] = coord[3 * o + c];
  }
  int s = 0;
  for (int a = 0; a < nion; ++a)
    while (s < ns && elem[order[a]] >= s) g->type_start[s++] = a;
  while (s <= ns) g->type_start[s++] = nion;
  for (int k = 0; k < ns; ++k) {
    std::snprintf(g->species[k].s, sizeof(g->species[k].s), "%s", names[k].c_str());
    std::snprintf(g->pseudo_file[k].s, sizeof(g->pseudo_file[k].s), "%s", pseudo[k].c_str());
  }
  return 0;
}

extern "C" int ofd_read_geometry(const char* path, ofd_geometry* g, char* err, int errlen) {
  try {
    return read_geometry(path, g, err, errlen);
  } catch (const std::exception& e) {
    if (g) ofd_free_geometry(g);
    set_err(err, errlen, e.what());
    return -1;
  }
}

extern "C" void ofd_free_geometry(ofd_geometry* g) {
  if (!g) return;
  std::free(g->type_start);
  std::free(g->frac);
  std::free(g->file_index);
  std::free(g->species);
  std::free(g->pseudo_file);
  g->type_start = nullptr;
  g->frac = nullptr;
  g->file_index = nullptr;
  g->species = nullptr;
  g->pseudo_file = nullptr;
}

static int read_recpot(const char* path, double* maxg_invA, double* values, int capacity, int* npts, char* err,
                       int errlen) {
  if (!path || !maxg_invA || !npts) {
    set_err(err, errlen, "null argument");
    return -1;
  }
  std::ifstream in(path);
  if (!in) {
    set_err(err, errlen, std::string("cannot open ") + path);
    return -1;
  }
  std::vector<std::string> lines;
  for (std::string ln; std::getline(in, ln);) lines.push_back(ln);
  size_t i = 0;
  while (i < lines.size() && lines[i].find("END COMMENT") == std::string::npos) ++i;
  if (i + 2 >= lines.size()) {
    set_err(err, errlen, "no END COMMENT block followed by the type and maxG lines");
    return -2;
  }
  // the type line ("3 5": reciprocal-space table) and maxG (1/A)
  auto tl = fields(lines[i + 1]);
  if (tl.empty() || tl[0] != "3") {
    set_err(err, errlen, "only reciprocal-space (type 3) tables are supported");
    return -2;
  }
  auto ml = fields(lines[i + 2]);
  if (ml.empty() || !to_double(ml[0], maxg_invA)) {
    set_err(err, errlen, "bad maxG line");
    return -2;
  }
  // numPoints = 3 x (lines after maxG, terminator included) - 3, then that
  // many values read list-directed; a short last data line leaves the
  // terminator as the last value, which reads as 0 (ReadIonFile.f90:858-895)
  const size_t first = i + 3;
  size_t nlines = 0;
  std::string last;
  for (size_t k = first; k < lines.size(); ++k) {
    auto f = fields(lines[k]);
    if (f.empty()) continue;
    ++nlines;
    last = f[0];
  }
  if (last.compare(0, 4, "1000") != 0) {
    set_err(err, errlen, "pseudopotential lacks the 1000 terminator");
    return -2;
  }
  const long long n = 3 * (long long)nlines - 3;
  std::vector<double> v;
  v.reserve((size_t)std::max(0LL, n));
  for (size_t k = first; k < lines.size() && (long long)v.size() < n; ++k)
    for (auto& t : fields(lines[k])) {
      if ((long long)v.size() >= n) break;
      double x;
      if (!to_double(t, &x)) {
        set_err(err, errlen, "not a number: " + t);
        return -2;
      }
      v.push_back(x);
    }
  if ((long long)v.size() < n) {
    set_err(err, errlen, "table shorter than its line count");
    return -2;
  }
  if (!v.empty() && v.back() == 1000.0) v.back() = 0.0;
  if (v.size() > (size_t)0x7fffffff) {
    set_err(err, errlen, "table too long");
    return -2;
  }
  *npts = (int)v.size();
  if (values) {
    if (capacity < (int)v.size()) {
      set_err(err, errlen, "capacity too small");
      return -3;
    }
    std::copy(v.begin(), v.end(), values);
  }
  return 0;
}

extern "C" int ofd_read_recpot(const char* path, double* maxg_invA, double* values, int capacity, int* npts,
                               char* err, int errlen) {
  try {
    return read_recpot(path, maxg_invA, values, capacity, npts, err, errlen);
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return -1;
  }
}

// One value as FORMAT(E26.20e2): 0.ddddddddddddddddddddE+xx (26 characters).
// The format has two exponent digits: magnitudes below 1e-99 are written as
// zero (the Fortran edit would fill the field with asterisks, which no reader
// accepts), and magnitudes of 1e99 and above are rejected.
static bool e26_20(double v, char* out) {
  if (!std::isfinite(v) || std::fabs(v) >= 1e99) return false;
  if (std::fabs(v) < 1e-99) v = 0.0;
  if (v == 0.0) {
    std::snprintf(out, 27, "%26s", "0.00000000000000000000E+00");
    return true;
  }
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.19E", std::fabs(v));   // d.ddd...E+xx, 20 significant digits
  const char* e = std::strchr(buf, 'E');
  const int ex = std::atoi(e + 1) + 1;
  char digits[32];
  int nd = 0;
  for (const char* p = buf; p < e; ++p)
    if (std::isdigit((unsigned char)*p)) digits[nd++] = *p;
  digits[nd] = '\0';
  if (v < 0)
    std::snprintf(out, 27, "-.%sE%c%02d", digits, ex < 0 ? '-' : '+', std::abs(ex));
  else
    std::snprintf(out, 27, "0.%sE%c%02d", digits, ex < 0 ? '-' : '+', std::abs(ex));
  return true;
}

extern "C" int ofd_write_density(const char* path, int n1, int n2, int n3, int nspin, const double* rho, char* err,
                                 int errlen) {
  if (!path || !rho || n1 <= 0 || n2 <= 0 || n3 <= 0 || nspin <= 0) {
    set_err(err, errlen, "bad argument");
    return -1;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_err(err, errlen, std::string("cannot create ") + path);
    return -1;
  }
  // 27-byte records: FORMAT(A14 I12.1) headers, FORMAT(E26.20e2) values
  // (Output.f90:684-700, 806-812); record 4 holds the spin count
  std::fprintf(f, "%-14s%12d\n", "x-dimension:", n1);
  std::fprintf(f, "%-14s%12d\n", "y-dimension:", n2);
  std::fprintf(f, "%-14s%12d\n", "z-dimension:", n3);
  std::fprintf(f, "%-14s%12d\n", "# of spins: ", nspin);
  const long long n = (long long)n1 * n2 * n3 * nspin;
  char rec[32];
  for (long long k = 0; k < n; ++k) {
    if (!e26_20(rho[k], rec)) {
      std::fclose(f);
      set_err(err, errlen, "value not representable in E26.20e2 at index " + std::to_string(k));
      return -2;
    }
    std::fprintf(f, "%s\n", rec);
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) set_err(err, errlen, "write failed");
  return ok ? 0 : -1;
}

static int read_density(const char* path, int n1, int n2, int n3, const double* cell, double num_electrons,
                        int repeat_x, int repeat_y, int repeat_z, double* rho, char* err, int errlen) {
  if (!path || !rho || n1 <= 0 || n2 <= 0 || n3 <= 0) {
    set_err(err, errlen, "bad argument");
    return -1;
  }
  std::ifstream in(path);
  if (!in) {
    set_err(err, errlen, std::string("cannot open ") + path);
    return -1;
  }
  std::string ln;
  int dims[4];
  for (int k = 0; k < 4; ++k) {
    if (!std::getline(in, ln) || ln.size() < 15) {
      set_err(err, errlen, "bad header record");
      return -2;
    }
    dims[k] = std::atoi(ln.substr(14).c_str());
  }
  const int xl = dims[0], yl = dims[1], zl = dims[2], fs = dims[3];
  if (xl <= 0 || yl <= 0 || zl <= 0) {
    set_err(err, errlen, "bad dimensions in the header");
    return -2;
  }
  if (fs != 1) {
    set_err(err, errlen, "only one spin channel is supported");
    return -2;
  }
  const long long nf = (long long)xl * yl * zl;
  if (nf > (1LL << 34)) {
    set_err(err, errlen, "header dimensions too large");
    return -2;
  }
  std::vector<double> d((size_t)nf);
  for (long long k = 0; k < nf; ++k) {
    if (!std::getline(in, ln)) {
      set_err(err, errlen, "file shorter than its dimensions");
      return -2;
    }
    auto f = fields(ln);
    if (f.empty() || !to_double(f[0], &d[(size_t)k])) {
      set_err(err, errlen, "bad value record");
      return -2;
    }
  }
  const int rx = repeat_x > 0 ? repeat_x : 1, ry = repeat_y > 0 ? repeat_y : 1, rz = repeat_z > 0 ? repeat_z : 1;
  auto at = [&](long long ix, long long iy, long long iz) { return d[(size_t)(ix + xl * (iy + yl * iz))]; };
  if ((long long)xl * rx == n1 && (long long)yl * ry == n2 && (long long)zl * rz == n3) {
    // same grid (with repeats): copy periodically (ReadIonFile.f90:1196-1210)
    for (int iz = 0; iz < n3; ++iz)
      for (int iy = 0; iy < n2; ++iy)
        for (int ix = 0; ix < n1; ++ix)
          rho[ix + (size_t)n1 * (iy + (size_t)n2 * iz)] = at(ix % xl, iy % yl, iz % zl);
  } else {
    // trilinear interpolation at the cell-centred positions of the new grid,
    // wrapping periodically (ReadIonFile.f90:1212-1288)
    auto frac_of = [](long long len, int rep, int n, int i) {
      double f = ((double)(len * rep) / (double)n * (1.0 + 2.0 * (double)i) - 1.0) * 0.5;
      f = f - (double)len * std::floor(f / (double)len);   // MODULO(f, len)
      // a tiny negative f (a rounded 0) wraps to len - tiny, which can round to
      // len itself: that is the point 0
      return f < (double)len ? f : 0.0;
    };
    for (int iz = 0; iz < n3; ++iz) {
      const double fz = frac_of(zl, rz, n3, iz);
      long long kz = (long long)fz;
      const double dz = fz - (double)kz;
      const long long sz = kz == zl - 1 ? zl : 0;
      for (int iy = 0; iy < n2; ++iy) {
        const double fy = frac_of(yl, ry, n2, iy);
        const long long ky = (long long)fy;
        const double dy = fy - (double)ky;
        const long long sy = ky == yl - 1 ? yl : 0;
        for (int ix = 0; ix < n1; ++ix) {
          const double fx = frac_of(xl, rx, n1, ix);
          const long long kx = (long long)fx;
          const double dx = fx - (double)kx;
          const long long sx = kx == xl - 1 ? xl : 0;
          auto plane = [&](long long z) {
            const double d1 = at(kx, ky, z), d2 = at(kx + 1 - sx, ky, z);
            const double d3 = at(kx, ky + 1 - sy, z), d4 = at(kx + 1 - sx, ky + 1 - sy, z);
            return (d1 * (1 - dx) + d2 * dx) * (1 - dy) + (d3 * (1 - dx) + d4 * dx) * dy;
          };
          const double p1 = plane(kz), p2 = plane(kz + 1 - sz);
          rho[ix + (size_t)n1 * (iy + (size_t)n2 * iz)] = p1 * (1 - dz) + p2 * dz;
        }
      }
    }
  }
  // scaled to num_electrons (ReadIonFile.f90:1290-1306); cell NULL or
  // num_electrons <= 0 leaves the values as read
  if (cell && num_electrons > 0) {
    double a[3][3];
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 3; ++c) a[c][j] = cell[3 * j + c];
    const double vol = std::fabs(a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                                 a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]));
    const long long n = (long long)n1 * n2 * n3;
    double integral = 0.0;
    for (long long k = 0; k < n; ++k) integral += rho[k];
    const double sc = num_electrons * (double)n1 * (double)n2 * (double)n3 / (integral * vol);
    for (long long k = 0; k < n; ++k) rho[k] *= sc;
  }
  return 0;
}

extern "C" int ofd_read_density(const char* path, int n1, int n2, int n3, const double* cell, double num_electrons,
                                int repeat_x, int repeat_y, int repeat_z, double* rho, char* err, int errlen) {
  try {
    return read_density(path, n1, n2, n3, cell, num_electrons, repeat_x, repeat_y, repeat_z, rho, err, errlen);
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return -1;
  }
}
