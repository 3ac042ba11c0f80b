// psx_io.cpp -- host-only file formats of the backend (no GPU needed):
//   CSV particles   ParticleData::readFromCSV / writeAsCSV (ParticleIO.cpp:8-144)
//   OBJ container   replaces PolygonContainer::readFromFile's assimp import
//                   (PolygonContainer.cpp:381-417): de-indexed triangles in face order
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "psx.h"
#include "psx_host.hpp"

namespace {

using psx::set_error;

// splitText (ParticleIO.cpp:8-19): getline on the delimiter, empty and "null" tokens dropped.
std::vector<std::string> split_tokens(const std::string& s, char delim) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, delim))
    if (!item.empty() && item != "null") out.push_back(item);
  return out;
}

// ParticleData::format(double, nChars) (ParticleIO.cpp:39-56): fixed notation with
// nChars - len(str(int(|v|))) - [v < 0] - 1 decimals (at least 0). The reference's (int) cast of
// |v| is evaluated as its x86-64 build does (cvttsd2si): NaN and |v| >= 2^31 give INT_MIN, whose 11
// characters leave no decimals, as any |v| >= 2^31 would.
std::string format_value(double value, int nchars) {
  std::ostringstream oss;
  const double a = std::abs(value);
  const int iv = (a < 2147483648.0) ? static_cast<int>(a) : std::numeric_limits<int>::min();
  int int_len = (int)std::to_string(iv).length();
  if (value < 0) int_len += 1;
  int dec = nchars - int_len - 1;
  if (dec < 0) dec = 0;
  oss << std::fixed << std::setprecision(dec) << value;
  return oss.str();
}

bool make_dirs(const std::string& dir) {
  if (dir.empty()) return true;
  std::string cur;
  for (size_t i = 0; i <= dir.size(); ++i) {
    if (i == dir.size() || dir[i] == '/') {
      if (!cur.empty() && cur != "/") {
        struct stat st;
        if (stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
          return false;
      }
    }
    if (i < dir.size()) cur.push_back(dir[i]);
  }
  return true;
}

template <class T>
T* copy_out(const std::vector<T>& v) {
  T* p = static_cast<T*>(std::malloc(std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

}  // namespace

extern "C" {

int psx_read_csv(const char* path, double density, const double* velocity3, int64_t* n,
                 double** pos, double** vel, double** radius, double** mass,
                 int64_t* n_skipped) {
  if (!path || !n || !pos || !vel || !radius || !mass)
    return set_error(PSX_ERR_INVALID, "NULL argument");
  std::ifstream file(path);
  if (!file.is_open()) return set_error(PSX_ERR_IO, std::string("Could not open file: ") + path);
  const double v0[3] = {velocity3 ? velocity3[0] : 0.0, velocity3 ? velocity3[1] : 0.0,
                        velocity3 ? velocity3[2] : 0.0};
  std::vector<double> P, V, R, M;
  std::string line;
  std::getline(file, line);  // header
  int64_t skipped = 0, lineno = 1;
  while (std::getline(file, line)) {
    ++lineno;
    const std::vector<std::string> tok = split_tokens(line, ',');
    if (tok.size() != 4) {  // ParticleIO.cpp:121-124: logged and skipped
      ++skipped;
      std::fprintf(stderr, "[psx] Invalid line: %s\n", line.c_str());
      continue;
    }
    double xyzr[4];
    for (int k = 0; k < 4; ++k) {
      try {
        xyzr[k] = std::stod(tok[(size_t)k]);
      } catch (const std::exception&) {
        return set_error(PSX_ERR_IO, std::string(path) + ":" + std::to_string(lineno) +
                                         ": not a number: '" + tok[(size_t)k] + "'");
      }
    }
    P.insert(P.end(), {xyzr[0], xyzr[1], xyzr[2]});
    V.insert(V.end(), {v0[0], v0[1], v0[2]});
    R.push_back(xyzr[3]);
    M.push_back(density * 4.0 / 3.0 * M_PI * std::pow(xyzr[3], 3));  // ParticleIO.cpp:131
  }
  *n = (int64_t)R.size();
  *pos = copy_out(P);
  *vel = copy_out(V);
  *radius = copy_out(R);
  *mass = copy_out(M);
  if (n_skipped) *n_skipped = skipped;
  if (!*pos || !*vel || !*radius || !*mass) return set_error(PSX_ERR_ALLOC, "host allocation failed");
  return PSX_OK;
}

int psx_write_csv(const char* path, int64_t n, const double* pos, const double* radius) {
  if (!path || n < 0 || (n > 0 && (!pos || !radius))) return set_error(PSX_ERR_INVALID, "bad argument");
  const std::string p(path);
  const size_t slash = p.find_last_of('/');
  if (slash != std::string::npos && !make_dirs(p.substr(0, slash)))
    return set_error(PSX_ERR_IO, "could not create directory for " + p);
  std::ofstream file(p);
  if (!file) return set_error(PSX_ERR_IO, "Could not open file " + p);
  file << "x,y,z,radius" << std::endl;
  for (int64_t i = 0; i < n; ++i)
    file << format_value(pos[3 * i], 10) << "," << format_value(pos[3 * i + 1], 10) << ","
         << format_value(pos[3 * i + 2], 10) << "," << format_value(radius[i], 10) << std::endl;
  file.close();
  if (!file) return set_error(PSX_ERR_IO, "write failed: " + p);
  return PSX_OK;
}

int psx_read_obj(const char* path, int64_t* n_triangles, double** vertices) {
  if (!path || !n_triangles || !vertices) return set_error(PSX_ERR_INVALID, "NULL argument");
  std::ifstream file(path);
  if (!file.is_open()) return set_error(PSX_ERR_IO, std::string("Could not open file: ") + path);
  std::vector<double> v, tri;
  std::string line;
  int64_t lineno = 0;
  while (std::getline(file, line)) {
    ++lineno;
    std::istringstream ss(line);
    std::string tag;
    if (!(ss >> tag)) continue;
    if (tag == "v") {
      double x, y, z;
      if (!(ss >> x >> y >> z))
        return set_error(PSX_ERR_IO, std::string(path) + ":" + std::to_string(lineno) + ": bad vertex");
      v.insert(v.end(), {x, y, z});
    } else if (tag == "f") {
      std::vector<int64_t> idx;
      std::string t;
      const int64_t nv = (int64_t)v.size() / 3;
      while (ss >> t) {
        char* end = nullptr;
        const long long k = std::strtoll(t.c_str(), &end, 10);
        if (end == t.c_str() || k == 0)
          return set_error(PSX_ERR_IO, std::string(path) + ":" + std::to_string(lineno) + ": bad face index");
        const int64_t i = k > 0 ? k - 1 : nv + k;
        if (i < 0 || i >= nv)
          return set_error(PSX_ERR_IO, std::string(path) + ":" + std::to_string(lineno) + ": face index out of range");
        idx.push_back(i);
      }
      if (idx.size() < 3)
        return set_error(PSX_ERR_IO, std::string(path) + ":" + std::to_string(lineno) + ": face with < 3 corners");
      for (size_t j = 1; j + 1 < idx.size(); ++j)  // fan triangulation of convex polygons
        for (int64_t c : {idx[0], idx[j], idx[j + 1]})
          tri.insert(tri.end(), {v[(size_t)(3 * c)], v[(size_t)(3 * c + 1)], v[(size_t)(3 * c + 2)]});
    }
  }
  if (tri.empty()) return set_error(PSX_ERR_IO, std::string(path) + ": no faces");
  *n_triangles = (int64_t)tri.size() / 9;
  *vertices = copy_out(tri);
  if (!*vertices) return set_error(PSX_ERR_ALLOC, "host allocation failed");
  return PSX_OK;
}

void psx_free(void* p) { std::free(p); }

}  // extern "C"
