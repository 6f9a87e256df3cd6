// Host-side setup for the fluid Newton step (not on the timed path):
//   * seeded synthetic pipe generator (hex grid -> 6 tets per hex),
//   * the block-CSR pattern the reference builds in lhsa (lhsa.cpp:153-380):
//     for every node, the sorted set of nodes that share an element with it,
//   * the FSILS views of it (inclusive rowPtr pairs and diagPtr, lhs.cpp:95-120),
//   * element colouring (no two elements of one colour share a node) and the
//     element -> block-slot map that replaces do_assem's binary search
//     (lhsa.cpp:123-134) on the device.
#include "svg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace svg {

// splitmix64: a seeded, platform-independent stream of 64-bit values.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next()
  {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [-1, 1)
  double sym() { return 2.0 * ((double)(next() >> 11) * (1.0 / 9007199254740992.0)) - 1.0; }
};

static double stenosis_scale(const svg_pipe_spec& sp, double z)
{
  if (sp.sten_depth <= 0.0 || sp.sten_len <= 0.0) return 1.0;
  if (z < sp.sten_z0 || z > sp.sten_z0 + sp.sten_len) return 1.0;
  const double pi = 3.14159265358979323846;
  return 1.0 - sp.sten_depth * 0.5 * (1.0 - std::cos(2.0 * pi * (z - sp.sten_z0) / sp.sten_len));
}

int pipe_counts(const svg_pipe_spec& sp, int* nNo, int* nEl)
{
  if (sp.n < 2 || (sp.n % 2) != 0 || sp.nz < 1) return set_error(SVG_ERR_ARG, "pipe: n must be even >= 2, nz >= 1");
  long long nn = (long long)(sp.n + 1) * (sp.n + 1) * (sp.nz + 1);
  long long ne = 6LL * sp.n * sp.n * sp.nz;
  if (nn > 0x7fffffffLL || ne > 0x7fffffffLL) return set_error(SVG_ERR_ARG, "pipe: too large for 32-bit indices");
  *nNo = (int)nn;
  *nEl = (int)ne;
  return SVG_OK;
}

int pipe_generate(const svg_pipe_spec& sp, double* x, int* IEN, int* flags)
{
  int nNo, nEl;
  int rc = pipe_counts(sp, &nNo, &nEl);
  if (rc) return rc;
  const int n = sp.n, nz = sp.nz;
  const double hs = 2.0 / n, hz = sp.length / nz;
  auto nid = [&](int i, int j, int k) { return i + (n + 1) * (j + (n + 1) * k); };
  SplitMix64 rng(sp.seed);
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= n; j++)
      for (int i = 0; i <= n; i++) {
        double s = -1.0 + hs * i, t = -1.0 + hs * j, z = hz * k;
        bool wall = (i == 0 || i == n || j == 0 || j == n);
        bool inlet = (k == 0), outlet = (k == nz);
        if (!wall && !inlet && !outlet) {
          s += sp.jitter * hs * rng.sym();
          t += sp.jitter * hs * rng.sym();
          z += sp.jitter * hz * rng.sym();
        }
        // Elliptical square -> disc map: (s,t) in [-1,1]^2 onto the unit disc.
        double X = s * std::sqrt(1.0 - 0.5 * t * t);
        double Y = t * std::sqrt(1.0 - 0.5 * s * s);
        double r = sp.radius * stenosis_scale(sp, z);
        int a = nid(i, j, k);
        x[3 * a + 0] = r * X;
        x[3 * a + 1] = r * Y;
        x[3 * a + 2] = z;
        if (flags) flags[a] = (inlet ? SVG_NODE_INLET : 0) | (wall ? SVG_NODE_WALL : 0) | (outlet ? SVG_NODE_OUTLET : 0);
      }
  // Six Kuhn tets per hex, mirrored per quadrant so that the face diagonal of
  // every cross-section corner cell passes through the square's corner.
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int e = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        bool fx = (i < n / 2), fy = (j < n / 2);
        auto corner = [&](int bx, int by, int bz) {
          int lx = fx ? 1 - bx : bx, ly = fy ? 1 - by : by;
          return nid(i + lx, j + ly, k + bz);
        };
        for (int p = 0; p < 6; p++) {
          int b[3] = {0, 0, 0};
          int v[4];
          v[0] = corner(0, 0, 0);
          for (int q = 0; q < 3; q++) {
            b[perm[p][q]] = 1;
            v[q + 1] = corner(b[0], b[1], b[2]);
          }
          // Positive orientation: det[x1-x0, x2-x0, x3-x0] > 0.
          const double* x0 = x + 3 * v[0];
          double d[3][3];
          for (int q = 0; q < 3; q++)
            for (int c = 0; c < 3; c++) d[q][c] = x[3 * v[q + 1] + c] - x0[c];
          double det = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                       d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                       d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
          if (det < 0) std::swap(v[2], v[3]);
          for (int a = 0; a < 4; a++) IEN[4 * e + a] = v[a];
          e++;
        }
      }
  return SVG_OK;
}

// ------------------------------------------------------------ bifurcation --
// A parent pipe (n x n cross-section on the disc, nzp layers) splitting into
// two daughters. Daughter d takes half of the parent's end cross-section
// (columns i <= n/2 for d = 0, i >= n/2 for d = 1), keeps that D-shaped
// section scaled towards its Murray radius r_d (r_0^3 = split r^3,
// r_1^3 = (1 - split) r^3), and turns by its angle about the y axis over the
// first Kt layers (d = 0 towards -x, d = 1 towards +x). The daughters' first
// layer is the parent's last; their hexes use the parent's per-quadrant Kuhn
// split by parent column, so the junction is conforming. Flags: inlet (parent
// z = 0), wall (outer surfaces and the carina line), outlet (daughter ends).
static void bif_layout(const svg_bif_spec& b, long long* nNo, long long* nEl)
{
  const long long n = b.n, half = n / 2;
  *nNo = (n + 1) * (n + 1) * (b.nzp + 1) + 2LL * (half + 1) * (n + 1) * b.nzd;
  *nEl = 6LL * n * n * b.nzp + 2LL * 6LL * half * n * b.nzd;
}

int bif_counts(const svg_bif_spec& b, int* nNo, int* nEl)
{
  if (b.n < 4 || (b.n % 4) != 0 || b.nzp < 1 || b.nzd < 2)
    return set_error(SVG_ERR_ARG, "bifurcation: n must be a multiple of 4 >= 4, nzp >= 1, nzd >= 2");
  if (!(b.split > 0.0 && b.split < 1.0)) return set_error(SVG_ERR_ARG, "bifurcation: split must be in (0, 1)");
  long long nn, ne;
  bif_layout(b, &nn, &ne);
  if (nn > 0x7fffffffLL || ne > 0x7fffffffLL) return set_error(SVG_ERR_ARG, "bifurcation: too large");
  *nNo = (int)nn;
  *nEl = (int)ne;
  return SVG_OK;
}

int bif_generate(const svg_bif_spec& b, double* x, int* IEN, int* flags)
{
  int nNo, nEl;
  int rc = bif_counts(b, &nNo, &nEl);
  if (rc) return rc;
  const int n = b.n, half = n / 2, nzp = b.nzp, nzd = b.nzd;
  const double pi = 3.14159265358979323846;
  const double hs = 2.0 / n, hzp = b.length_parent / nzp, hzd = b.length_daughter / nzd;
  const int Kt = std::max(3, nzd / 3);
  SplitMix64 rng(b.seed);
  const int np = (n + 1) * (n + 1) * (nzp + 1), nd = (half + 1) * (n + 1) * nzd;
  auto pid = [&](int i, int j, int k) { return i + (n + 1) * (j + (n + 1) * k); };
  // daughter d, local column ii in [0, half], layer k in [0, nzd] (k = 0: the parent's last layer)
  auto did = [&](int d, int ii, int j, int k) {
    if (k == 0) return pid(ii + d * half, j, nzp);
    return np + d * nd + ii + (half + 1) * (j + (n + 1) * (k - 1));
  };
  auto disc = [&](int i, int j, double& X, double& Y) {
    const double s = -1.0 + hs * i, t = -1.0 + hs * j;
    X = b.radius * s * std::sqrt(1.0 - 0.5 * t * t);
    Y = b.radius * t * std::sqrt(1.0 - 0.5 * s * s);
  };
  for (int k = 0; k <= nzp; k++)
    for (int j = 0; j <= n; j++)
      for (int i = 0; i <= n; i++) {
        const bool wall = (i == 0 || i == n || j == 0 || j == n) || (k == nzp && i == half);
        const bool inlet = (k == 0);
        double X, Y, z = hzp * k;
        disc(i, j, X, Y);
        if (!wall && !inlet) {
          X += b.jitter * hs * b.radius * rng.sym();
          Y += b.jitter * hs * b.radius * rng.sym();
          if (k < nzp) z += b.jitter * hzp * rng.sym();
        }
        const int a = pid(i, j, k);
        x[3 * a + 0] = X;
        x[3 * a + 1] = Y;
        x[3 * a + 2] = z;
        if (flags) flags[a] = (inlet ? SVG_NODE_INLET : 0) | (wall ? SVG_NODE_WALL : 0);
      }
  for (int d = 0; d < 2; d++) {
    const double sg = d == 0 ? -1.0 : 1.0;
    const double rd = std::cbrt(d == 0 ? b.split : 1.0 - b.split) * b.radius;
    const double rho_d = rd / b.radius;
    const double th_d = d == 0 ? b.angle0 : b.angle1;
    const double xc = sg * 4.0 * b.radius / (3.0 * pi);  // centroid of the half disc
    double o[3] = {xc, 0.0, b.length_parent};
    for (int k = 1; k <= nzd; k++) {
      const double thm = th_d * std::min(1.0, (k - 0.5) / Kt), th = th_d * std::min(1.0, (double)k / Kt);
      o[0] += hzd * sg * std::sin(thm);
      o[2] += hzd * std::cos(thm);
      const double rho = 1.0 + (rho_d - 1.0) * std::min(1.0, (double)k / Kt);
      const double u[3] = {std::cos(th), 0.0, -sg * std::sin(th)};
      const double dir[3] = {sg * std::sin(th), 0.0, std::cos(th)};
      for (int j = 0; j <= n; j++)
        for (int ii = 0; ii <= half; ii++) {
          const bool wall = (ii == 0 || ii == half || j == 0 || j == n);
          const bool outlet = (k == nzd);
          double X, Y;
          disc(ii + d * half, j, X, Y);
          double qx = rho * (X - xc), qy = rho * Y, qz = 0.0;
          if (!wall && !outlet) {
            const double hl = b.jitter * hs * rd;
            qx += hl * rng.sym();
            qy += hl * rng.sym();
            qz += b.jitter * hzd * rng.sym();
          }
          const int a = did(d, ii, j, k);
          for (int c = 0; c < 3; c++) x[3 * a + c] = o[c] + qx * u[c] + qz * dir[c] + (c == 1 ? qy : 0.0);
          if (flags) flags[a] = (wall ? SVG_NODE_WALL : 0) | (outlet ? SVG_NODE_OUTLET : 0);
        }
    }
  }
  // Six Kuhn tets per hex, mirrored per parent quadrant (see pipe_generate).
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int e = 0;
  auto emit = [&](auto&& corner) {
    for (int p = 0; p < 6; p++) {
      int bb[3] = {0, 0, 0};
      int v[4];
      v[0] = corner(0, 0, 0);
      for (int q = 0; q < 3; q++) {
        bb[perm[p][q]] = 1;
        v[q + 1] = corner(bb[0], bb[1], bb[2]);
      }
      const double* x0 = x + 3 * v[0];
      double dd[3][3];
      for (int q = 0; q < 3; q++)
        for (int c = 0; c < 3; c++) dd[q][c] = x[3 * v[q + 1] + c] - x0[c];
      const double det = dd[0][0] * (dd[1][1] * dd[2][2] - dd[1][2] * dd[2][1]) -
                         dd[0][1] * (dd[1][0] * dd[2][2] - dd[1][2] * dd[2][0]) +
                         dd[0][2] * (dd[1][0] * dd[2][1] - dd[1][1] * dd[2][0]);
      if (det < 0) std::swap(v[2], v[3]);
      for (int a = 0; a < 4; a++) IEN[4 * e + a] = v[a];
      e++;
    }
  };
  for (int k = 0; k < nzp; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const bool fx = (i < half), fy = (j < half);
        emit([&](int bx, int by, int bz) {
          return pid(i + (fx ? 1 - bx : bx), j + (fy ? 1 - by : by), k + bz);
        });
      }
  for (int d = 0; d < 2; d++)
    for (int k = 0; k < nzd; k++)
      for (int j = 0; j < n; j++)
        for (int ii = 0; ii < half; ii++) {
          const bool fx = (ii + d * half < half), fy = (j < half);
          emit([&](int bx, int by, int bz) {
            return did(d, ii + (fx ? 1 - bx : bx), j + (fy ? 1 - by : by), k + bz);
          });
        }
  return SVG_OK;
}

// Node -> element incidence in CSR form.
static void node_elements(int nNo, int nEl, int eNoN, const int* IEN, std::vector<int>& ptr, std::vector<int>& lst)
{
  ptr.assign(nNo + 1, 0);
  for (long long q = 0; q < (long long)nEl * eNoN; q++) ptr[IEN[q] + 1]++;
  for (int a = 0; a < nNo; a++) ptr[a + 1] += ptr[a];
  lst.resize(ptr[nNo]);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < nEl; e++)
    for (int a = 0; a < eNoN; a++) lst[fill[IEN[(size_t)eNoN * e + a]]++] = e;
}

int csr_pattern(int nNo, int nEl, int eNoN, const int* IEN, std::vector<int>& rowPtr, std::vector<int>& colPtr)
{
  for (long long q = 0; q < (long long)nEl * eNoN; q++)
    if (IEN[q] < 0 || IEN[q] >= nNo) return set_error(SVG_ERR_ARG, "csr_pattern: IEN entry out of range");
  std::vector<int> nptr, nlst;
  node_elements(nNo, nEl, eNoN, IEN, nptr, nlst);
  rowPtr.assign(nNo + 1, 0);
  colPtr.clear();
  std::vector<int> buf;
  for (int a = 0; a < nNo; a++) {
    buf.clear();
    for (int q = nptr[a]; q < nptr[a + 1]; q++) {
      int e = nlst[q];
      for (int b = 0; b < eNoN; b++) buf.push_back(IEN[(size_t)eNoN * e + b]);
    }
    buf.push_back(a);
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    colPtr.insert(colPtr.end(), buf.begin(), buf.end());
    rowPtr[a + 1] = (int)colPtr.size();
    if (colPtr.size() > 0x7fffffffULL / 16) return set_error(SVG_ERR_ARG, "csr_pattern: 16*nnz overflows int");
  }
  return SVG_OK;
}

int color_elements(int nNo, int nEl, int eNoN, const int* IEN, std::vector<int>& color, int* nColors)
{
  // Greedy colouring in element order: smallest colour unused by any earlier
  // element sharing a node. Per-node bit masks of used colours (up to 256).
  const int W = 4;  // 4 x 64 bits
  std::vector<uint64_t> used((size_t)nNo * W, 0);
  color.assign(nEl, -1);
  int nc = 0;
  for (int e = 0; e < nEl; e++) {
    uint64_t m[W] = {0, 0, 0, 0};
    for (int a = 0; a < eNoN; a++) {
      const uint64_t* u = &used[(size_t)W * IEN[(size_t)eNoN * e + a]];
      for (int w = 0; w < W; w++) m[w] |= u[w];
    }
    int c = -1;
    for (int w = 0; w < W && c < 0; w++)
      if (~m[w]) c = 64 * w + __builtin_ctzll(~m[w]);
    if (c < 0) return set_error(SVG_ERR_ARG, "color_elements: more than 256 colours needed");
    color[e] = c;
    nc = std::max(nc, c + 1);
    for (int a = 0; a < eNoN; a++) used[(size_t)W * IEN[(size_t)eNoN * e + a] + c / 64] |= (1ull << (c % 64));
  }
  *nColors = nc;
  return SVG_OK;
}

int slot_map(int nEl, int eNoN, const int* IEN, const int* rowPtr, const int* colPtr, std::vector<int>& slots)
{
  slots.resize((size_t)nEl * eNoN * eNoN);
  for (int e = 0; e < nEl; e++)
    for (int a = 0; a < eNoN; a++) {
      int r = IEN[(size_t)eNoN * e + a];
      const int* beg = colPtr + rowPtr[r];
      const int* end = colPtr + rowPtr[r + 1];
      for (int b = 0; b < eNoN; b++) {
        int c = IEN[(size_t)eNoN * e + b];
        const int* it = std::lower_bound(beg, end, c);
        if (it == end || *it != c) return set_error(SVG_ERR_ARG, "slot_map: pattern does not contain an element pair");
        slots[((size_t)e * eNoN + a) * eNoN + b] = (int)(it - colPtr);
      }
    }
  return SVG_OK;
}

}  // namespace svg

extern "C" {

int svg_pipe_counts(const svg_pipe_spec* sp, int* nNo, int* nEl)
{
  if (!sp || !nNo || !nEl) return svg::set_error(SVG_ERR_ARG, "svg_pipe_counts: null argument");
  return svg::pipe_counts(*sp, nNo, nEl);
}

int svg_pipe_generate(const svg_pipe_spec* sp, double* x, int* IEN, int* flags)
{
  if (!sp || !x || !IEN) return svg::set_error(SVG_ERR_ARG, "svg_pipe_generate: null argument");
  return svg::pipe_generate(*sp, x, IEN, flags);
}

int svg_bif_counts(const svg_bif_spec* sp, int* nNo, int* nEl)
{
  if (!sp || !nNo || !nEl) return svg::set_error(SVG_ERR_ARG, "svg_bif_counts: null argument");
  return svg::bif_counts(*sp, nNo, nEl);
}

int svg_bif_generate(const svg_bif_spec* sp, double* x, int* IEN, int* flags)
{
  if (!sp || !x || !IEN) return svg::set_error(SVG_ERR_ARG, "svg_bif_generate: null argument");
  return svg::bif_generate(*sp, x, IEN, flags);
}

int svg_csr_pattern_count(int nNo, int nEl, int eNoN, const int* IEN, int* nnz)
{
  std::vector<int> r, c;
  int rc = svg::csr_pattern(nNo, nEl, eNoN, IEN, r, c);
  if (rc) return rc;
  *nnz = (int)c.size();
  return SVG_OK;
}

int svg_csr_pattern(int nNo, int nEl, int eNoN, const int* IEN, int* rowPtr, int* colPtr)
{
  std::vector<int> r, c;
  int rc = svg::csr_pattern(nNo, nEl, eNoN, IEN, r, c);
  if (rc) return rc;
  std::memcpy(rowPtr, r.data(), sizeof(int) * r.size());
  std::memcpy(colPtr, c.data(), sizeof(int) * c.size());
  return SVG_OK;
}

int svg_color_elements(int nNo, int nEl, int eNoN, const int* IEN, int* color, int* nColors)
{
  std::vector<int> c;
  int rc = svg::color_elements(nNo, nEl, eNoN, IEN, c, nColors);
  if (rc) return rc;
  std::memcpy(color, c.data(), sizeof(int) * c.size());
  return SVG_OK;
}

}  // extern "C"
