// alv_oracle.cpp -- CPU oracle for the SPH solver core.
//
// TEST INFRASTRUCTURE ONLY (see alv_oracle.h).  A plain C++17/OpenMP
// restatement of the reference algorithm, templated on the working precision
// TF (float or double).  Each function cites the reference file:line it
// follows.  Floating-point expressions keep the reference's operation order;
// the build uses -ffp-contract=off so that no multiply-add is fused.
//
// Deliberate, documented choices where the reference is non-deterministic:
//  * Cell membership is a counting sort (ascending particle id inside a cell)
//    instead of atomic insertion (runner.hpp:1840-1867), so neighbour order is
//    canonical: 27 cells in the reference order, ascending id inside a cell.
//    Cells are not capped at max_num_particles_per_cell; the overflow flag
//    (grid_anomaly[1]) is still raised.
//  * Convergence and rigid-force sums accumulate in FP64, serially in particle
//    order (the reference uses FP32 thrust::reduce, runner.hpp:4646-4663).
//  * rsqrt in the boundary normal is 1/sqrt (correctly rounded).

#include "alv_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

template <class T>
static void volume_from_distance_impl(double h, const double* dmin, const double* dmax,
                                      const unsigned* res, const double* cs, double sign,
                                      const double* dist, unsigned begin, unsigned end,
                                      double* out);

namespace orc {

// Planted mistakes (orc_plant): a test switches one on to show that its
// check fails on a plausible error in the oracle.  0 = the reference's
// algorithm.
enum Plant {
  kPlantNone = 0,
  kPlantFactorNoGradSum = 1,     // DFSPH factor without |sum_j grad p_j|^2
  kPlantFactorBoundarySign = 2,  // boundary gradient subtracted in the factor
  kPlantDivWarmClamp = 3,        // divergence warm start clamped at -0.25
  kPlantDensWarmClamp = 4,       // density warm start clamped at -5e-4
  kPlantDivFinish = 5,           // stored kappa_v = kappa_v dt^2
  kPlantDensFinish = 6,          // stored kappa = kappa dt
  kPlantIiOmega = 7,             // IISPH relaxation weight 1/2
  kPlantIiAiiNoDji = 8,          // IISPH a_ii without the d_ji term
  kPlantIiDiiDensity = 9,        // IISPH d_ii with 1/rho instead of 1/rho^2
  kPlantCflRadius = 10,          // CFL length scale r instead of 2r
};
static int g_plant = kPlantNone;

template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
struct V4 {
  T x, y, z, w;
};
template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
inline V3<T> operator-(V3<T> a) {
  return {-a.x, -a.y, -a.z};
}
template <class T>
inline V3<T> operator*(T s, V3<T> a) {
  return {s * a.x, s * a.y, s * a.z};
}
template <class T>
inline V3<T> operator*(V3<T> a, T s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class T>
inline V3<T> operator*(V3<T> a, V3<T> b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <class T>
inline V3<T>& operator+=(V3<T>& a, V3<T> b) {
  a = a + b;
  return a;
}
template <class T>
inline V3<T>& operator-=(V3<T>& a, V3<T> b) {
  a = a - b;
  return a;
}
template <class T>
inline T dot(V3<T> a, V3<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
inline T length_sqr(V3<T> a) {
  return dot(a, a);
}

// ---------------------------------------------------------------------------
// Peirce quadrature constants (reference dg/peirce_quadrature.hpp:10-13 and
// the M = 14 table).  Regenerated here: cos(phi_j), b_j are the 30-point
// Gauss-Legendre rule; r_k = sqrt((x_k + 1) / 2) and c_k = w_k / (4 sqrt 2)
// with (x_k, w_k) the 15-point Gauss-Jacobi(alpha = 0, beta = 1/2) rule, so
// that sum_k c_k g(r_k) integrates g(r) r^2 over [0, 1].  Both rules come
// from the Golub-Welsch eigenproblem solved by implicit-shift QL.
// ---------------------------------------------------------------------------
constexpr int kNumPhi = 30, kNumTheta = 60, kNumR = 15;

static void golub_welsch(int n, std::vector<long double>& diag,
                         std::vector<long double>& off, long double mu0,
                         std::vector<long double>& nodes,
                         std::vector<long double>& weights) {
  std::vector<long double> d = diag, e(n, 0.0L), z(n, 0.0L);
  for (int i = 0; i + 1 < n; ++i) e[i] = off[i];
  z[0] = 1.0L;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        long double dd = fabsl(d[m]) + fabsl(d[m + 1]);
        if (fabsl(e[m]) <= 1e-30L * dd + 1e-300L) break;
      }
      if (m != l) {
        if (++iter > 200) break;
        long double g = (d[l + 1] - d[l]) / (2.0L * e[l]);
        long double r = hypotl(g, 1.0L);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabsl(r) : -fabsl(r)));
        long double s = 1.0L, c = 1.0L, p = 0.0L;
        int i;
        bool underflow = false;
        for (i = m - 1; i >= l; --i) {
          long double f = s * e[i];
          long double b = c * e[i];
          r = hypotl(f, g);
          e[i + 1] = r;
          if (r == 0.0L) {
            d[i + 1] -= p;
            e[m] = 0.0L;
            underflow = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0L * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          long double zf = z[i + 1];
          z[i + 1] = s * z[i] + c * zf;
          z[i] = c * z[i] - s * zf;
        }
        if (underflow) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0L;
      }
    } while (m != l);
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  nodes.resize(n);
  weights.resize(n);
  for (int i = 0; i < n; ++i) {
    nodes[i] = d[idx[i]];
    weights[i] = mu0 * z[idx[i]] * z[idx[i]];
  }
}

struct PeirceTable {
  double cosphi[kNumPhi], sinphi[kNumPhi], b[kNumPhi], r[kNumR], c[kNumR];
  PeirceTable() {
    {
      int n = kNumPhi;
      std::vector<long double> diag(n, 0.0L), off(n - 1), x, w;
      for (int k = 1; k < n; ++k) {
        long double kk = k;
        off[k - 1] = sqrtl(kk * kk / (4.0L * kk * kk - 1.0L));
      }
      golub_welsch(n, diag, off, 2.0L, x, w);
      for (int j = 0; j < n; ++j) {
        cosphi[j] = static_cast<double>(x[j]);
        sinphi[j] = static_cast<double>(sqrtl(1.0L - x[j] * x[j]));
        b[j] = static_cast<double>(w[j]);
      }
    }
    {
      int n = kNumR;
      const long double al = 0.0L, be = 0.5L;
      std::vector<long double> diag(n), off(n - 1), x, w;
      for (int k = 0; k < n; ++k) {
        long double s = 2.0L * k + al + be;
        diag[k] = (be * be - al * al) / (s * (s + 2.0L));
      }
      for (int k = 1; k < n; ++k) {
        long double s = 2.0L * k + al + be;
        long double num = 4.0L * k * (k + al) * (k + be) * (k + al + be);
        long double den = s * s * (s + 1.0L) * (s - 1.0L);
        off[k - 1] = sqrtl(num / den);
      }
      // mu0 = integral of (1+x)^(1/2) over [-1, 1] = 2^(5/2) / 3
      long double mu0 = powl(2.0L, 2.5L) / 3.0L;
      golub_welsch(n, diag, off, mu0, x, w);
      for (int k = 0; k < n; ++k) {
        r[k] = static_cast<double>(sqrtl((x[k] + 1.0L) / 2.0L));
        c[k] = static_cast<double>(w[k] / (4.0L * sqrtl(2.0L)));
      }
    }
  }
};
static const PeirceTable& peirce() {
  static PeirceTable t;
  return t;
}

// ---------------------------------------------------------------------------
// Constants (reference constants.hpp:92-137)
// ---------------------------------------------------------------------------
template <class T>
struct Consts {
  T h, h2, k, l, w0;
  T radius, vol, mass, rho0;
  T visc, bvisc, eps;
  V3<T> g;
  int wrap_mask;
  T wrap_len[3], wrap_min[3], wrap_max[3];
  int res[3], off[3];
  unsigned kmax, cell_cap;
  T cosphi[kNumPhi], sinphi[kNumPhi], B[kNumPhi], R[kNumR], C[kNumR];
  T pi;

  void set(const orc_params& p) {
    pi = static_cast<T>(
        3.14159265358979323846264338327950288419716939937510582097494459230781640628620899862803482534211706798214808651e+00L);
    // set_kernel_radius (constants.hpp:100-120)
    T hh = static_cast<T>(p.kernel_radius);
    T hh2 = hh * hh;
    T hh3 = hh2 * hh;
    T hh4 = hh2 * hh2;
    h = hh;
    h2 = hh2;
    k = static_cast<T>(8.0) / hh3 / pi;
    l = static_cast<T>(48.0) / hh4 / pi;
    w0 = k;
    // set_particle_attr (constants.hpp:122-128)
    T m = static_cast<T>(p.particle_mass);
    T rho = static_cast<T>(p.density0);
    vol = m / rho;
    radius = static_cast<T>(p.particle_radius);
    mass = m;
    rho0 = rho;
    visc = static_cast<T>(p.viscosity);
    bvisc = static_cast<T>(p.boundary_viscosity);
    eps = static_cast<T>(p.dfsph_factor_epsilon);
    g = {static_cast<T>(p.gravity[0]), static_cast<T>(p.gravity[1]),
         static_cast<T>(p.gravity[2])};
    wrap_mask = 0;
    for (int a = 0; a < 3; ++a) {
      // set_wrap_length (constants.hpp:130-136)
      T L = static_cast<T>(p.wrap_length[a]);
      wrap_len[a] = L;
      wrap_min[a] = L * static_cast<T>(-0.5);
      wrap_max[a] = L * static_cast<T>(0.5);
      if (p.wrap_length[a] > 0) wrap_mask |= 1 << a;
      res[a] = p.grid_res[a];
      off[a] = p.grid_offset[a];
    }
    kmax = p.max_num_neighbors_per_particle;
    cell_cap = p.max_num_particles_per_cell;
    const PeirceTable& t = peirce();
    for (int j = 0; j < kNumPhi; ++j) {
      cosphi[j] = static_cast<T>(t.cosphi[j]);
      sinphi[j] = static_cast<T>(t.sinphi[j]);
      B[j] = static_cast<T>(t.b[j]);
    }
    for (int q = 0; q < kNumR; ++q) {
      R[q] = static_cast<T>(t.r[q]);
      C[q] = static_cast<T>(t.c[q]);
    }
  }
};

// ---------------------------------------------------------------------------
// Cubic spline kernel (runner.hpp:596-638)
// ---------------------------------------------------------------------------
template <class T>
inline T dist_cubic_kernel(T r, const Consts<T>& c) {
  T q = std::fabs(r) / c.h;
  T conj = 1 - q;
  T result = 0;
  if (q <= static_cast<T>(0.5))
    result = (6 * q - 6) * q * q + 1;
  else if (q <= 1)
    result = 2 * conj * conj * conj;
  return result * c.k;
}
template <class T>
inline T d2_cubic_kernel(T r2, const Consts<T>& c) {
  T q2 = r2 / c.h2;
  T q = std::sqrt(r2) / c.h;
  T conj = 1 - q;
  T result = 0;
  if (q <= static_cast<T>(0.5))
    result = (6 * q - 6) * q2 + 1;
  else if (q <= 1)
    result = 2 * conj * conj * conj;
  return result * c.k;
}
template <class T>
inline T displacement_cubic_kernel(V3<T> d, const Consts<T>& c) {
  return d2_cubic_kernel(length_sqr(d), c);
}
template <class T>
inline V3<T> displacement_cubic_kernel_grad(V3<T> d, const Consts<T>& c) {
  T rl2 = length_sqr(d);
  T rl = std::sqrt(rl2);
  T q = rl / c.h;
  T scale = (q <= static_cast<T>(0.5) ? (q * 3 - 2) / c.h
                                      : (q <= 1 ? (1 - q) * (q - 1) / rl : 0));
  return (c.l * scale) * d;
}

// ---------------------------------------------------------------------------
// Quaternion helpers (runner.hpp:120-258).  q = (x, y, z, w).
// ---------------------------------------------------------------------------
template <class T>
inline V3<T> rotate(V3<T> v, V4<T> q) {
  T xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  T xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
  T wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  V3<T> r;
  r.x = (1 - 2 * (yy + zz)) * v.x + 2 * (xy - wz) * v.y + 2 * (xz + wy) * v.z;
  r.y = 2 * (xy + wz) * v.x + (1 - 2 * (xx + zz)) * v.y + 2 * (yz - wx) * v.z;
  r.z = 2 * (xz - wy) * v.x + 2 * (yz + wx) * v.y + (1 - 2 * (xx + yy)) * v.z;
  return r;
}
template <class T>
inline V4<T> conjugate(V4<T> q) {
  return {-q.x, -q.y, -q.z, q.w};
}
template <class T>
inline void orthogonal_vectors(V3<T> n, V3<T>& t1, V3<T>& t2) {
  // get_orthogonal_vectors (runner.hpp:120-131)
  V3<T> v{1, 0, 0};
  if (std::fabs(dot(v, n)) > static_cast<T>(0.999)) {
    v.x = 0;
    v.y = 1;
  }
  t1 = cross(n, v);
  t2 = cross(n, t1);
  t1 = t1 * (1 / std::sqrt(length_sqr(t1)));
  t2 = t2 * (1 / std::sqrt(length_sqr(t2)));
}

// ---------------------------------------------------------------------------
// Analytic SDFs (dg/box_distance.hpp:23-39, dg/box_shell_distance.hpp:36-59,
// dg/sphere_distance.hpp:18-21).  Device variants are restated.
// ---------------------------------------------------------------------------
template <class T>
struct Sdf {
  int type;
  V3<T> half, half_outer;
  T outset, radius;
  T max_distance;
  V3<T> aabb_min, aabb_max;

  static T box_dist(V3<T> x, V3<T> half) {
    V3<T> diff{std::fabs(x.x) - half.x, std::fabs(x.y) - half.y,
               std::fabs(x.z) - half.z};
    V3<T> cl{std::max(diff.x, T(0)), std::max(diff.y, T(0)),
             std::max(diff.z, T(0))};
    return std::sqrt(length_sqr(cl)) +
           std::min(std::max(diff.x, std::max(diff.y, diff.z)), T(0));
  }
  T signed_distance(V3<T> x) const {
    if (type == ORC_BOX) return box_dist(x, half) - outset;
    if (type == ORC_BOX_SHELL) {
      T inner = box_dist(x, half) - outset;
      T outer = box_dist(x, half_outer) - outset;
      return std::max(outer, -inner);
    }
    return std::sqrt(length_sqr(x)) - radius;
  }
  V3<T> gradient(V3<T> x, T scale) const {
    if (type == ORC_SPHERE) return x;
    const T kEps = static_cast<T>(0.00390625);
    T step = scale * kEps;
    return {signed_distance(x + V3<T>{step, 0, 0}) -
                signed_distance(x - V3<T>{step, 0, 0}),
            signed_distance(x + V3<T>{0, step, 0}) -
                signed_distance(x - V3<T>{0, step, 0}),
            signed_distance(x + V3<T>{0, 0, step}) -
                signed_distance(x - V3<T>{0, 0, step})};
  }
  void set(const orc_boundary_desc& d) {
    type = d.type;
    if (type == ORC_BOX) {
      V3<T> w{T(d.p[0]), T(d.p[1]), T(d.p[2])};
      half = w * T(0.5);
      outset = T(d.p[3]);
      aabb_min = {w.x * T(-0.5) - outset, w.y * T(-0.5) - outset,
                  w.z * T(-0.5) - outset};
      aabb_max = {half.x + outset, half.y + outset, half.z + outset};
      max_distance = std::sqrt(length_sqr(half)) + outset;
    } else if (type == ORC_BOX_SHELL) {
      V3<T> w{T(d.p[0]), T(d.p[1]), T(d.p[2])};
      T th = T(d.p[3]);
      half = w * T(0.5);
      half_outer = {half.x + th, half.y + th, half.z + th};
      outset = T(d.p[4]);
      aabb_min = {w.x * T(-0.5) - th - outset, w.y * T(-0.5) - th - outset,
                  w.z * T(-0.5) - th - outset};
      aabb_max = {half.x + th + outset, half.y + th + outset,
                  half.z + th + outset};
      max_distance = std::sqrt(length_sqr(half_outer)) + outset;
    } else if (type == ORC_MAP) {
      // the distance comes from the boundary's map; only its box is kept
      aabb_min = {T(d.aabb_min[0]), T(d.aabb_min[1]), T(d.aabb_min[2])};
      aabb_max = {T(d.aabb_max[0]), T(d.aabb_max[1]), T(d.aabb_max[2])};
      V3<T> far{std::max(std::fabs(aabb_min.x), std::fabs(aabb_max.x)),
                std::max(std::fabs(aabb_min.y), std::fabs(aabb_max.y)),
                std::max(std::fabs(aabb_min.z), std::fabs(aabb_max.z))};
      max_distance = std::sqrt(length_sqr(far));
    } else {
      radius = T(d.p[0]);
      outset = 0;
      aabb_min = {-radius, -radius, -radius};
      aabb_max = {radius, radius, radius};
      max_distance = radius;
    }
  }
};

// ---------------------------------------------------------------------------
// 32-node cubic Lagrange volume-map machinery (runner.hpp:1131-1594)
// ---------------------------------------------------------------------------
template <class T>
struct GridGeom {
  V3<T> dmin, dmax, cs;
  unsigned res[3];
  unsigned num_nodes;
};

template <class T>
V3<T> node_position(const GridGeom<T>& g, unsigned l) {
  // index_to_node_position (runner.hpp:1131-1178)
  const unsigned rx = g.res[0], ry = g.res[1], rz = g.res[2];
  unsigned nv = (rx + 1) * (ry + 1) * (rz + 1);
  unsigned ne_x = rx * (ry + 1) * (rz + 1);
  unsigned ne_y = (rx + 1) * ry * (rz + 1);
  V3<T> x;
  if (l < nv) {
    unsigned temp = l % ((ry + 1) * (rx + 1));
    V3<T> idx{T(temp % (rx + 1)), T(temp / (rx + 1)),
              T(l / ((ry + 1) * (rx + 1)))};
    x = g.dmin + g.cs * idx;
  } else if (l < nv + 2 * ne_x) {
    l -= nv;
    unsigned e = l / 2;
    unsigned temp = e % ((ry + 1) * rx);
    V3<T> idx{T(temp % rx), T(temp / rx), T(e / ((ry + 1) * rx))};
    x = g.dmin + g.cs * idx;
    x.x += (static_cast<T>(1) + (l % 2)) / static_cast<T>(3) * g.cs.x;
  } else if (l < nv + 2 * (ne_x + ne_y)) {
    l -= (nv + 2 * ne_x);
    unsigned e = l / 2;
    unsigned temp = e % ((rz + 1) * ry);
    V3<T> idx{T(e / ((rz + 1) * ry)), T(temp % ry), T(temp / ry)};
    x = g.dmin + g.cs * idx;
    x.y += (static_cast<T>(1) + (l % 2)) / static_cast<T>(3) * g.cs.y;
  } else {
    l -= (nv + 2 * (ne_x + ne_y));
    unsigned e = l / 2;
    unsigned temp = e % ((rx + 1) * rz);
    V3<T> idx{T(temp / rz), T(e / ((rx + 1) * rz)), T(temp % rz)};
    x = g.dmin + g.cs * idx;
    x.z += (static_cast<T>(1) + (l % 2)) / static_cast<T>(3) * g.cs.z;
  }
  return x;
}

template <class T>
bool resolve(const GridGeom<T>& g, V3<T> x, int ipos[3], V3<T>& inner) {
  // resolve (runner.hpp:1180-1195)
  V3<T> inv{1 / g.cs.x, 1 / g.cs.y, 1 / g.cs.z};
  if (x.x >= g.dmin.x && x.y >= g.dmin.y && x.z >= g.dmin.z &&
      g.dmax.x >= x.x && g.dmax.y >= x.y && g.dmax.z >= x.z) {
    V3<T> f = (x - g.dmin) * inv;
    ipos[0] = std::min(static_cast<int>(f.x), static_cast<int>(g.res[0]) - 1);
    ipos[1] = std::min(static_cast<int>(f.y), static_cast<int>(g.res[1]) - 1);
    ipos[2] = std::min(static_cast<int>(f.z), static_cast<int>(g.res[2]) - 1);
    V3<T> fi{T(ipos[0]), T(ipos[1]), T(ipos[2])};
    V3<T> sd_min = g.dmin + g.cs * fi;
    V3<T> t = (x - sd_min) * inv;
    inner = {2 * t.x - 1, 2 * t.y - 1, 2 * t.z - 1};
    return true;
  }
  return false;
}

inline void cell_nodes(const unsigned res[3], const int ip[3], unsigned* cell) {
  // get_cells (runner.hpp:1197-1257): 8 corners, then 2 nodes per edge for
  // the x-, y- and z-directed edges of the cell.
  const unsigned rx = res[0], ry = res[1], rz = res[2];
  const unsigned i = ip[0], j = ip[1], k = ip[2];
  const unsigned sx = rx + 1, sxy = (rx + 1) * (ry + 1);
  for (unsigned c = 0; c < 8; ++c) {
    unsigned di = c & 1u, dj = (c >> 1) & 1u, dk = (c >> 2) & 1u;
    cell[c] = sxy * (k + dk) + sx * (j + dj) + (i + di);
  }
  unsigned base = (rx + 1) * (ry + 1) * (rz + 1);
  // x-directed edges: (dj, dk) in order (0,0), (0,1), (1,0), (1,1)
  for (unsigned e = 0; e < 4; ++e) {
    unsigned dj = e >> 1, dk = e & 1u;
    unsigned n = base + 2 * (rx * (ry + 1) * (k + dk) + rx * (j + dj) + i);
    cell[8 + 2 * e] = n;
    cell[9 + 2 * e] = n + 1;
  }
  base += 2 * (rx * (ry + 1) * (rz + 1));
  // y-directed edges: (di, dk) in order (0,0), (1,0), (0,1), (1,1)
  for (unsigned e = 0; e < 4; ++e) {
    unsigned di = e & 1u, dk = e >> 1;
    unsigned n = base + 2 * (ry * (rz + 1) * (i + di) + ry * (k + dk) + j);
    cell[16 + 2 * e] = n;
    cell[17 + 2 * e] = n + 1;
  }
  base += 2 * ((rx + 1) * ry * (rz + 1));
  // z-directed edges: (di, dj) in order (0,0), (0,1), (1,0), (1,1)
  for (unsigned e = 0; e < 4; ++e) {
    unsigned di = e >> 1, dj = e & 1u;
    unsigned n = base + 2 * (rz * (rx + 1) * (j + dj) + rz * (i + di) + k);
    cell[24 + 2 * e] = n;
    cell[25 + 2 * e] = n + 1;
  }
}

// Serendipity-type cubic element on [-1,1]^3 (Discregrid's 32-node
// CubicLagrangeDiscreteGrid, restated by runner.hpp:1259-1584).
// Corner node (a,b,c) in {-1,1}^3:
//   N = (1/64) (1+a x)(1+b y)(1+c z) (9 (x^2+y^2+z^2) - 19)
// Edge node along x at x = s/3 (s = -1, +1), y = b, z = c:
//   N = (9/64) (1-x^2) (1+3 s x) (1+b y)(1+c z)    (y, z edges likewise)
template <class T>
void shape_functions(V3<T> xi, T* N, T* dNx, T* dNy, T* dNz) {
  const T x = xi.x, y = xi.y, z = xi.z;
  const T k64 = 1 / static_cast<T>(64);
  const T k964 = 9 / static_cast<T>(64);
  const T F = k64 * (9 * (x * x + y * y + z * z) - 19);
  for (int c = 0; c < 8; ++c) {
    T a = (c & 1) ? T(1) : T(-1);
    T b = (c & 2) ? T(1) : T(-1);
    T cc = (c & 4) ? T(1) : T(-1);
    T A = 1 + a * x, Bf = 1 + b * y, Cf = 1 + cc * z;
    N[c] = F * (A * Bf) * Cf;
    if (dNx) {
      dNx[c] = (a * F + A * (k64 * (18 * x))) * (Bf * Cf);
      dNy[c] = (b * F + Bf * (k64 * (18 * y))) * (A * Cf);
      dNz[c] = (cc * F + Cf * (k64 * (18 * z))) * (A * Bf);
    }
  }
  // edge families: axis 0 (x-directed), 1 (y-directed), 2 (z-directed)
  for (int axis = 0; axis < 3; ++axis) {
    T u = axis == 0 ? x : (axis == 1 ? y : z);
    for (int e = 0; e < 4; ++e) {
      // the two transverse signs, in the node order of cell_nodes()
      T p, q;  // signs for the (first, second) transverse coordinate
      if (axis == 0) {  // transverse (y, z); order (y-,z-),(y-,z+),(y+,z-),(y+,z+)
        p = (e >> 1) ? T(1) : T(-1);
        q = (e & 1) ? T(1) : T(-1);
      } else if (axis == 1) {  // transverse (x, z); order x fastest
        p = (e & 1) ? T(1) : T(-1);
        q = (e >> 1) ? T(1) : T(-1);
      } else {  // transverse (x, y); order y fastest
        p = (e >> 1) ? T(1) : T(-1);
        q = (e & 1) ? T(1) : T(-1);
      }
      T v1, v2;  // transverse coordinates
      if (axis == 0) {
        v1 = y;
        v2 = z;
      } else if (axis == 1) {
        v1 = x;
        v2 = z;
      } else {
        v1 = x;
        v2 = y;
      }
      T P = 1 + p * v1, Q = 1 + q * v2;
      T one_m_u2 = 1 - u * u;
      for (int s2 = 0; s2 < 2; ++s2) {
        T s = s2 ? T(1) : T(-1);
        T L = 1 + 3 * s * u;
        int n = 8 + axis * 8 + 2 * e + s2;
        T G = k964 * (one_m_u2 * L);
        N[n] = G * (P * Q);
        if (dNx) {
          T dG = k964 * (-2 * u * L + one_m_u2 * (3 * s));
          T d_along = dG * (P * Q);
          T d_v1 = G * (p * Q);
          T d_v2 = G * (P * q);
          T* dU = axis == 0 ? dNx : (axis == 1 ? dNy : dNz);
          T* dV1 = axis == 0 ? dNy : dNx;
          T* dV2 = axis == 2 ? dNy : dNz;
          dU[n] = d_along;
          dV1[n] = d_v1;
          dV2[n] = d_v2;
        }
      }
    }
  }
}

template <class T, class Nodes>
T interpolate(const Nodes& nodes, const unsigned* cells, const T* N) {
  // interpolate (runner.hpp:1586-1594)
  T phi = 0;
  for (int j = 0; j < 32; ++j) phi += nodes[cells[j]] * N[j];
  return phi;
}

// interpolate_and_derive (runner.hpp:1773-1791): the value and the
// gradient, scaled from the local coordinate to world units by 2 / cell size
template <class T, class Nodes>
T interpolate_and_derive(const Nodes& nodes, const unsigned* cells, const T* N,
                         const T* dNx, const T* dNy, const T* dNz, V3<T> cs,
                         V3<T>& grad) {
  T phi = 0;
  V3<T> g{0, 0, 0};
  for (int j = 0; j < 32; ++j) {
    T c = nodes[cells[j]];
    phi += c * N[j];
    g.x += c * dNx[j];
    g.y += c * dNy[j];
    g.z += c * dNz[j];
  }
  grad = {g.x * (T(2) / cs.x), g.y * (T(2) / cs.y), g.z * (T(2) / cs.z)};
  return phi;
}

// ---------------------------------------------------------------------------
// Boundary description (Pile, pile.hpp:40-78, 285-352)
// ---------------------------------------------------------------------------
template <class T>
struct Boundary {
  Sdf<T> sdf;
  T sign;
  GridGeom<T> geom;
  std::vector<T> volume;
  T mass;
  V3<T> inertia;
  V3<T> x, v, omega, force, torque, a;
  V4<T> q;
  // rigid contacts (Pile::add collision_mesh, restitution, friction)
  T restitution = 1, friction = 0;
  std::vector<V3<T>> coll;  // collision vertices in the body frame
  std::vector<T> distance;  // ORC_MAP: distance map nodes
  bool is_map() const { return sdf.type == ORC_MAP; }
  // Distance (before the sign) and its gradient at body-frame point `local`
  // from the distance map; false outside the map.
  bool map_distance(V3<T> local, T& d, V3<T>& grad) const {
    int ip[3];
    V3<T> inner;
    if (!resolve(geom, local, ip, inner)) return false;
    T N[32], dNx[32], dNy[32], dNz[32];
    unsigned cells[32];
    shape_functions<T>(inner, N, dNx, dNy, dNz);
    cell_nodes(geom.res, ip, cells);
    d = interpolate_and_derive(distance, cells, N, dNx, dNy, dNz, geom.cs, grad);
    return true;
  }
  // Collision distance and normal of a body-frame point
  // (collision_find_dist_normal, runner.hpp:1794-1825, for a map; the
  // analytic variant otherwise); the normal is unit when dist < 0.
  T contact_dist_normal(V3<T> local, T tolerance, V3<T>& n) const {
    T dist;
    if (is_map()) {
      T d;
      V3<T> g;
      dist = sign;
      n = {0, 0, 0};
      if (map_distance(local, d, g)) {
        dist = sign * d - tolerance;
        n = g * sign;
      }
    } else {
      dist = sdf.signed_distance(local) * sign - tolerance;
      n = sdf.gradient(local, static_cast<T>(1e-3)) * sign;
    }
    if (dist < 0) n = n * (1 / std::sqrt(length_sqr(n)));
    return dist;
  }
};

// Contact<float3, float> (contact.hpp:7-24)
template <class T>
struct Contact {
  unsigned i, j;
  V3<T> cp_i, cp_j, n, t, iiwi_diag, iiwi_off_diag, iiwj_diag, iiwj_off_diag;
  T friction, nkninv, pmax, goalu, impulse_sum;
};

// ---------------------------------------------------------------------------
// Solver state
// ---------------------------------------------------------------------------
template <class T>
struct Solver {
  int kind;  // 0 DF, 1 II, 2 ISPH (SolverI)
  unsigned max_n, n;
  Consts<T> c;
  orc_settings set;
  orc_report rep;
  T t = 0, dt = 0;
  std::vector<Boundary<T>> bnd;
  V3<T> pile_gravity{0, 0, 0};
  T contact_tolerance = 0;
  unsigned max_num_contacts = 0;
  unsigned num_found = 0;
  std::vector<Contact<T>> contacts;

  std::vector<V3<T>> x, v, a;
  std::vector<T> density, dfsph_factor, kappa, kappa_v, density_adv, cfl_v2;
  // IISPH
  std::vector<T> pressure, last_pressure, aii, adv_density, density_err;
  std::vector<V3<T>> dii, dij_pj, pressure_accel;
  // boundary records [B][N]
  std::vector<V4<T>> boundary, boundary_kernel;
  std::vector<V3<T>> force, torque;
  // grid and neighbours
  std::vector<unsigned> cell_count, cell_start, cell_ids;
  std::vector<V3<T>> grid_x;  // positions at the last grid build (the pid grid's entries)
  std::vector<unsigned> nb_id, num_neighbors;
  std::vector<V3<T>> nb_x;
  unsigned anomaly[3] = {0, 0, 0};
  // Particle-sampled boundaries (Pile with VolumeMethod::pellets,
  // pile.hpp:16,217-290): the collision vertices of every body are its
  // pellets; pellet k of the concatenation occupies particle slot max_n + k
  // of x, v, pressure, last_pressure and the neighbour lists
  // (solver.hpp:30-85).
  bool pellet_mode = false;
  unsigned max_p = 0, np = 0;
  std::vector<unsigned> pellet_body;  // pellet_id_to_rigid_id (pile.hpp:624-636)
  std::vector<unsigned> pnb_id, num_pnb;  // pellet neighbours [max_n + max_p][kmax]
  std::vector<V3<T>> pnb_x;
  std::vector<std::pair<unsigned, unsigned>> ident;  // hint_identical_sequence
  struct Kin {
    V3<T> x, v, omega;
    V4<T> q;
  };
  std::vector<Kin> kin0;  // body kinematics as copied to the device at the step start

  size_t nt() const { return size_t(max_n) + max_p; }
  unsigned num_cells() const {
    return unsigned(c.res[0]) * unsigned(c.res[1]) * unsigned(c.res[2]);
  }

  void alloc() {
    x.assign(nt(), {0, 0, 0});
    v.assign(nt(), {0, 0, 0});
    a.assign(max_n, {0, 0, 0});
    density.assign(max_n, 0);
    dfsph_factor.assign(max_n, 0);
    kappa.assign(max_n, 0);
    kappa_v.assign(max_n, 0);
    density_adv.assign(max_n, 0);
    cfl_v2.assign(max_n, 0);
    pressure.assign(nt(), 0);
    last_pressure.assign(nt(), 0);
    aii.assign(max_n, 0);
    adv_density.assign(max_n, 0);
    density_err.assign(max_n, 0);
    dii.assign(max_n, {0, 0, 0});
    dij_pj.assign(max_n, {0, 0, 0});
    pressure_accel.assign(max_n, {0, 0, 0});
    num_neighbors.assign(nt(), 0);
    nb_id.assign(nt() * c.kmax, 0);
    nb_x.assign(nt() * c.kmax, {0, 0, 0});
    num_pnb.assign(nt(), 0);
    pnb_id.assign(pellet_mode ? nt() * c.kmax : 0, 0);
    pnb_x.assign(pellet_mode ? nt() * c.kmax : 0, {0, 0, 0});
    cell_count.assign(num_cells(), 0);
    cell_start.assign(num_cells() + 1, 0);
    realloc_boundary();
  }
  void realloc_boundary() {
    size_t nb = bnd.size() * size_t(max_n);
    boundary.assign(nb, {0, 0, 0, 0});
    boundary_kernel.assign(nb, {0, 0, 0, 0});
    force.assign(nb, {0, 0, 0});
    torque.assign(nb, {0, 0, 0});
  }

  // ---- neighbour grid -----------------------------------------------------
  void ipos_of(V3<T> p, int ip[3]) const {
    // get_ipos (runner.hpp:1827-1831)
    ip[0] = static_cast<int>(std::floor(p.x / c.h)) - c.off[0];
    ip[1] = static_cast<int>(std::floor(p.y / c.h)) - c.off[1];
    ip[2] = static_cast<int>(std::floor(p.z / c.h)) - c.off[2];
  }
  bool within_grid(const int ip[3]) const {
    return 0 <= ip[0] && ip[0] < c.res[0] && 0 <= ip[1] && ip[1] < c.res[1] &&
           0 <= ip[2] && ip[2] < c.res[2];
  }
  unsigned linear(const int ip[3]) const {
    // Variable<3,U> row-major, z fastest (variable.hpp:286-290)
    return (unsigned(ip[0]) * c.res[1] + unsigned(ip[1])) * c.res[2] +
           unsigned(ip[2]);
  }
  V3<T> wrap_disp(V3<T> d) const {
    // wrap_y (runner.hpp:699-707), per enabled axis
    T* comp[3] = {&d.x, &d.y, &d.z};
    for (int ax = 0; ax < 3; ++ax) {
      if (!(c.wrap_mask & (1 << ax))) continue;
      if (*comp[ax] >= c.wrap_max[ax])
        *comp[ax] -= c.wrap_len[ax];
      else if (*comp[ax] < c.wrap_min[ax])
        *comp[ax] += c.wrap_len[ax];
    }
    return d;
  }

  void update_particle_grid() {
    // update_particle_grid (runner.hpp:1840-1867) as a counting sort
    unsigned C = num_cells();
    std::fill(cell_count.begin(), cell_count.end(), 0u);
    anomaly[0] = anomaly[1] = anomaly[2] = 0;
    // fluid particles, then (pellet mode) the pellets: launch_update_particle_grid
    // over [0, n) and [max_n, max_n + np) (solver.hpp:229-240)
    std::vector<unsigned> ids(n);
    for (unsigned i = 0; i < n; ++i) ids[i] = i;
    if (pellet_mode)
      for (unsigned k = 0; k < np; ++k) ids.push_back(max_n + k);
    std::vector<unsigned> key(nt(), C);
    for (unsigned i : ids) {
      int ip[3];
      ipos_of(x[i], ip);
      if (within_grid(ip)) {
        key[i] = linear(ip);
        cell_count[key[i]]++;
      } else {
        key[i] = C;
        anomaly[0] = 1;
      }
    }
    unsigned acc = 0;
    for (unsigned k = 0; k < C; ++k) {
      cell_start[k] = acc;
      acc += cell_count[k];
      if (cell_count[k] > c.cell_cap) anomaly[1] = 1;
    }
    cell_start[C] = acc;
    cell_ids.assign(acc, 0);
    std::vector<unsigned> fill(C, 0);
    for (unsigned i : ids) {
      if (key[i] == C) continue;
      cell_ids[cell_start[key[i]] + fill[key[i]]++] = i;
    }
    grid_x.assign(x.begin(), pellet_mode ? x.end() : x.begin() + n);
  }

  void make_neighbor_list() {
    // make_neighbor_list<wrap> (runner.hpp:1869-1911)
    const unsigned K = c.kmax;
    int flag2 = 0;
#pragma omp parallel for schedule(static) reduction(| : flag2)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      V3<T> xi = x[i];
      int ip[3];
      ipos_of(xi, ip);
      unsigned cnt = 0;
      for (int t = 0; t < 27; ++t) {
        int np[3] = {ip[0] + t / 9 - 1, ip[1] + (t / 3) % 3 - 1,
                     ip[2] + t % 3 - 1};
        for (int ax = 0; ax < 3; ++ax) {
          if (!(c.wrap_mask & (1 << ax))) continue;
          // wrap_ipos_y (runner.hpp:689-697), per enabled axis
          if (np[ax] >= c.res[ax])
            np[ax] -= c.res[ax];
          else if (np[ax] < 0)
            np[ax] += c.res[ax];
        }
        if (!within_grid(np)) continue;
        unsigned cell = linear(np);
        for (unsigned s = cell_start[cell]; s < cell_start[cell + 1]; ++s) {
          unsigned j = cell_ids[s];
          V3<T> xixj = wrap_disp(xi - x[j]);
          if (j != i && length_sqr(xixj) < c.h2) {
            size_t slot = size_t(i) * K + std::min(K - 1, cnt);
            nb_id[slot] = j;
            nb_x[slot] = xixj;
            cnt += 1;
          }
        }
      }
      num_neighbors[i] = std::min(K, cnt);
      if (cnt > K) flag2 = 1;
    }
    if (flag2) anomaly[2] = 1;
  }

  // ---- boundaries ---------------------------------------------------------
  void compute_particle_boundary(unsigned b) {
    // compute_particle_boundary_analytic<0> (runner.hpp:1982-2002) with
    // compute_volume_and_boundary_x_analytic (runner.hpp:1738-1771)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      size_t o = size_t(b) * max_n + i;
      point_boundary(b, x[i], boundary[o], boundary_kernel[o]);
    }
  }
  void point_boundary(unsigned b, V3<T> xi, V4<T>& rec, V4<T>& ker) const {
    const Boundary<T>& B = bnd[b];
    if (B.is_map()) {
      // compute_volume_and_boundary_x (runner.hpp:1696-1736) with
      // compute_particle_boundary (runner.hpp:2005-2025): outside the map
      // the record is the particle itself with zero volume
      V3<T> local = rotate(xi - B.x, conjugate(B.q));
      V3<T> bx = xi, xi_bxj{0, 0, 0};
      T d = 0, vol = 0;
      int ip[3];
      V3<T> inner;
      if (resolve(B.geom, local, ip, inner)) {
        T N[32], dNx[32], dNy[32], dNz[32];
        unsigned cells[32];
        shape_functions<T>(inner, N, dNx, dNy, dNz);
        cell_nodes(B.geom.res, ip, cells);
        V3<T> g;
        d = interpolate_and_derive(B.distance, cells, N, dNx, dNy, dNz, B.geom.cs, g) *
            B.sign;
        V3<T> normal = rotate(g * B.sign, B.q);
        T nl2 = length_sqr(normal);
        normal = normal * (nl2 > 0 ? 1 / std::sqrt(nl2) : T(0));
        xi_bxj = d * normal;
        bx = xi - xi_bxj;
        vol = interpolate(B.volume, cells, N);
      }
      V3<T> kg = displacement_cubic_kernel_grad(xi_bxj, c);
      rec = {bx.x, bx.y, bx.z, vol};
      ker = {vol * kg.x, vol * kg.y, vol * kg.z, vol * dist_cubic_kernel(d, c)};
      return;
    }
    {
      V3<T> local = rotate(xi - B.x, conjugate(B.q));
      T d = B.sdf.signed_distance(local) * B.sign;
      V3<T> normal = rotate(B.sdf.gradient(local, c.h) * B.sign, B.q);
      T nl2 = length_sqr(normal);
      normal = normal * (nl2 > 0 ? 1 / std::sqrt(nl2) : T(0));
      V3<T> xi_bxj = d * normal;
      V3<T> bx = xi - xi_bxj;
      T vol = 0;
      int ip[3];
      V3<T> inner;
      if (resolve(B.geom, local, ip, inner)) {
        T N[32];
        unsigned cells[32];
        shape_functions<T>(inner, N, nullptr, nullptr, nullptr);
        cell_nodes(B.geom.res, ip, cells);
        vol = interpolate(B.volume, cells, N);
      }
      V3<T> kg = displacement_cubic_kernel_grad(xi_bxj, c);
      rec = {bx.x, bx.y, bx.z, vol};
      ker = {vol * kg.x, vol * kg.y, vol * kg.z, vol * dist_cubic_kernel(d, c)};
    }
  }

  // ---- sampling (runner.hpp:1869-1911 on sample points, 2100-2120,
  // 4173-4324) ---------------------------------------------------------------
  // Sample neighbour lists: ids [ns][K], displacements [ns][K] from the grid
  // positions; a particle whose id equals the sample index is skipped.
  void sample_neighbors(const V3<T>* sx, unsigned ns, unsigned* ids, V3<T>* xixj,
                        unsigned* counts) {
    const unsigned K = c.kmax;
    for (unsigned i = 0; i < ns; ++i) {
      int ip[3];
      ipos_of(sx[i], ip);
      unsigned cnt = 0;
      for (int t = 0; t < 27; ++t) {
        int np[3] = {ip[0] + t / 9 - 1, ip[1] + (t / 3) % 3 - 1, ip[2] + t % 3 - 1};
        for (int ax = 0; ax < 3; ++ax) {
          if (!(c.wrap_mask & (1 << ax))) continue;
          if (np[ax] >= c.res[ax])
            np[ax] -= c.res[ax];
          else if (np[ax] < 0)
            np[ax] += c.res[ax];
        }
        if (!within_grid(np)) continue;
        unsigned cell = linear(np);
        for (unsigned s = cell_start[cell]; s < cell_start[cell + 1]; ++s) {
          unsigned j = cell_ids[s];
          V3<T> d = wrap_disp(sx[i] - grid_x[j]);
          if (j != i && length_sqr(d) < c.h2) {
            size_t slot = size_t(i) * K + std::min(K - 1, cnt);
            ids[slot] = j;
            xixj[slot] = d;
            cnt += 1;
          }
        }
      }
      counts[i] = std::min(K, cnt);
      if (cnt > K) anomaly[2] = 1;
    }
  }
  // sample_fluid: sum_j m / rho_j W q_j (q has `dim` components)
  void sample_fluid(const unsigned* ids, const V3<T>* xixj, const unsigned* counts, unsigned ns,
                    const T* q, unsigned dim, T* out) const {
    const unsigned K = c.kmax;
    for (unsigned i = 0; i < ns; ++i) {
      T acc[3] = {0, 0, 0};
      for (unsigned k = 0; k < counts[i]; ++k) {
        size_t o = size_t(i) * K + k;
        unsigned j = ids[o];
        T w = c.mass / density[j] * displacement_cubic_kernel(xixj[o], c);
        for (unsigned a = 0; a < dim; ++a) acc[a] += w * q[size_t(j) * dim + a];
      }
      for (unsigned a = 0; a < dim; ++a) out[size_t(i) * dim + a] = acc[a];
    }
  }
  // sample_fluid_density, or sample_position_density with boundary kernels
  void sample_density(const V3<T>* xixj, const unsigned* counts, unsigned ns, const V4<T>* ker,
                      unsigned nb, T* out) const {
    const unsigned K = c.kmax;
    for (unsigned i = 0; i < ns; ++i) {
      T dens = 0;
      for (unsigned k = 0; k < counts[i]; ++k)
        dens += c.vol * displacement_cubic_kernel(xixj[size_t(i) * K + k], c);
      for (unsigned b = 0; b < nb; ++b) dens += ker[size_t(b) * ns + i].w;
      out[i] = dens * c.rho0;
    }
  }
  // compute_sample_velocity, or compute_sample_vorticity when sv is given
  void sample_velocity(const unsigned* ids, const V3<T>* xixj, const unsigned* counts,
                       unsigned ns, const V3<T>* sv, const V4<T>* rec, const V4<T>* ker,
                       unsigned nb, const V3<T>* rx, const V3<T>* rv, const V3<T>* rw,
                       V3<T>* out) const {
    const unsigned K = c.kmax;
    for (unsigned i = 0; i < ns; ++i) {
      V3<T> acc{0, 0, 0};
      V3<T> vi = sv ? sv[i] : V3<T>{0, 0, 0};
      for (unsigned k = 0; k < counts[i]; ++k) {
        size_t o = size_t(i) * K + k;
        unsigned j = ids[o];
        if (sv)
          acc += c.mass / density[j] * cross(v[j] - vi, displacement_cubic_kernel_grad(xixj[o], c));
        else
          acc += (c.mass / density[j] * displacement_cubic_kernel(xixj[o], c)) * v[j];
      }
      for (unsigned b = 0; b < nb; ++b) {
        V4<T> r = rec[size_t(b) * ns + i], kk = ker[size_t(b) * ns + i];
        V3<T> bx{r.x, r.y, r.z};
        V3<T> velj = cross(rw[b], bx - rx[b]) + rv[b];
        if (sv)
          acc += cross(velj - vi, V3<T>{kk.x, kk.y, kk.z});
        else
          acc += velj * kk.w;
      }
      out[i] = acc;
    }
  }
  void compute_all_boundaries() {
    for (unsigned b = 0; b < bnd.size(); ++b) compute_particle_boundary(b);
  }
  V4<T> brec(unsigned b, unsigned i) const {
    return boundary[size_t(b) * max_n + i];
  }
  V4<T> bker(unsigned b, unsigned i) const {
    return boundary_kernel[size_t(b) * max_n + i];
  }
  V3<T> rigid_vel(unsigned b, V3<T> bx) const {
    return cross(bnd[b].omega, bx - bnd[b].x) + bnd[b].v;
  }
  void add_force(unsigned b, unsigned i, V3<T> bx, V3<T> f) {
    size_t o = size_t(b) * max_n + i;
    force[o] += f;
    torque[o] += cross(bx - bnd[b].x, f);
  }

  // ---- density and DFSPH kernels --------------------------------------------
  void compute_density() {
    // compute_density (runner.hpp:2052-2072)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      T dens = c.vol * c.w0;
      for (unsigned s = 0; s < num_neighbors[i]; ++s)
        dens += c.vol * displacement_cubic_kernel(nb_x[size_t(i) * c.kmax + s], c);
      for (unsigned b = 0; b < bnd.size(); ++b) dens += bker(b, i).w;
      density[i] = dens * c.rho0;
    }
  }
  void compute_dfsph_factor() {
    // compute_dfsph_factor (runner.hpp:3124-3154)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      T sum_grad_p_k = 0;
      V3<T> grad_p_i{0, 0, 0};
      for (unsigned s = 0; s < num_neighbors[i]; ++s) {
        V3<T> grad_p_j =
            (-c.vol) * displacement_cubic_kernel_grad(nb_x[size_t(i) * c.kmax + s], c);
        sum_grad_p_k += length_sqr(grad_p_j);
        grad_p_i -= grad_p_j;
      }
      for (unsigned b = 0; b < bnd.size(); ++b) {
        V4<T> k = bker(b, i);
        if (g_plant == kPlantFactorBoundarySign)
          grad_p_i -= V3<T>{k.x, k.y, k.z};
        else
          grad_p_i += V3<T>{k.x, k.y, k.z};
      }
      if (g_plant != kPlantFactorNoGradSum) sum_grad_p_k += length_sqr(grad_p_i);
      dfsph_factor[i] = sum_grad_p_k > c.eps ? -1 / sum_grad_p_k : T(0);
    }
  }
  T density_change(unsigned i) const {
    // compute_density_change (runner.hpp:3192-3228)
    T density_adv_i = 0;
    V3<T> vi = v[i];
    for (unsigned s = 0; s < num_neighbors[i]; ++s) {
      size_t o = size_t(i) * c.kmax + s;
      density_adv_i += c.vol * dot(vi - v[nb_id[o]],
                                   displacement_cubic_kernel_grad(nb_x[o], c));
    }
    for (unsigned b = 0; b < bnd.size(); ++b) {
      V4<T> r = brec(b, i), k = bker(b, i);
      V3<T> velj = rigid_vel(b, {r.x, r.y, r.z});
      density_adv_i += dot(vi - velj, V3<T>{k.x, k.y, k.z});
    }
    density_adv_i = std::max(density_adv_i, static_cast<T>(0));
    if (num_neighbors[i] < 20) density_adv_i = 0;
    return density_adv_i;
  }
  T density_adv_of(unsigned i, T dt_) const {
    // compute_density_adv (runner.hpp:3493-3523)
    T delta = 0;
    V3<T> vi = v[i];
    for (unsigned s = 0; s < num_neighbors[i]; ++s) {
      size_t o = size_t(i) * c.kmax + s;
      delta += c.vol * dot(vi - v[nb_id[o]],
                           displacement_cubic_kernel_grad(nb_x[o], c));
    }
    for (unsigned b = 0; b < bnd.size(); ++b) {
      V4<T> r = brec(b, i), k = bker(b, i);
      V3<T> velj = rigid_vel(b, {r.x, r.y, r.z});
      delta += dot(vi - velj, V3<T>{k.x, k.y, k.z});
    }
    return std::max(density[i] / c.rho0 + dt_ * delta, static_cast<T>(1));
  }
  // Shared body of warm_start_divergence_solve_1, divergence_solve_iteration,
  // warm_start_pressure_solve1 and pressure_solve_iteration
  // (runner.hpp:3252-3295, 3316-3363, 3452-3491, 3543-3588).
  // kj(j) gives neighbour j's stiffness; sign_div selects the divergence
  // form (dv -= dt*ksum*(-V gradW)) vs the pressure form
  // (dv += dt*ksum*V*gradW).
  template <class KJ>
  V3<T> kappa_dv(unsigned i, T k_i, T dt_, bool divergence_form, KJ kj) {
    V3<T> dv{0, 0, 0};
    for (unsigned s = 0; s < num_neighbors[i]; ++s) {
      size_t o = size_t(i) * c.kmax + s;
      T k_sum = k_i + kj(nb_id[o]);
      if (std::fabs(k_sum) > c.eps) {
        if (divergence_form) {
          V3<T> grad_p_j = (-c.vol) * displacement_cubic_kernel_grad(nb_x[o], c);
          dv -= (dt_ * k_sum) * grad_p_j;
        } else {
          dv += ((dt_ * k_sum) * c.vol) * displacement_cubic_kernel_grad(nb_x[o], c);
        }
      }
    }
    if (std::fabs(k_i) > c.eps) {
      for (unsigned b = 0; b < bnd.size(); ++b) {
        V4<T> r = brec(b, i), k = bker(b, i);
        V3<T> grad_wvol{k.x, k.y, k.z};
        V3<T> acc = k_i * (-grad_wvol);
        V3<T> f = c.mass * acc;
        dv -= dt_ * acc;
        add_force(b, i, {r.x, r.y, r.z}, f);
      }
    }
    return dv;
  }

  void compute_viscosity() {
    // compute_viscosity (runner.hpp:2150-2249)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      V3<T> vi = v[i], xi = x[i];
      V3<T> da{0, 0, 0};
      for (unsigned s = 0; s < num_neighbors[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = nb_id[o];
        V3<T> xixj = nb_x[o];
        da += (c.visc * c.mass / density[j] *
               dot(xixj, displacement_cubic_kernel_grad(xixj, c)) /
               (length_sqr(xixj) + static_cast<T>(0.01) * c.h2)) *
              (vi - v[j]);
      }
      for (unsigned b = 0; b < bnd.size(); ++b) {
        V4<T> rec = brec(b, i);
        V3<T> bx{rec.x, rec.y, rec.z};
        V3<T> r_x = bnd[b].x, r_v = bnd[b].v, r_omega = bnd[b].omega;
        V3<T> normal = bx - xi;
        T nl2 = length_sqr(normal);
        if (nl2 > 0) {
          normal = normal * (1 / std::sqrt(nl2));
          V3<T> t1, t2;
          orthogonal_vectors(normal, t1, t2);
          T dist = c.h - std::sqrt(nl2);
          V3<T> xs[4] = {bx - t1 * dist, bx + t1 * dist, bx - t2 * dist,
                         bx + t2 * dist};
          T vol = static_cast<T>(0.25) * rec.w;
          V3<T> fsum{0, 0, 0}, tsum{0, 0, 0};
          V3<T> as[4];
          for (int p = 0; p < 4; ++p) {
            V3<T> xix = xi - xs[p];
            V3<T> gradW = displacement_cubic_kernel_grad(xix, c);
            V3<T> vp = cross(r_omega, xs[p] - r_x) + r_v;
            as[p] = (c.bvisc * vol * dot(xix, gradW) /
                     (length_sqr(xix) + static_cast<T>(0.01) * c.h2)) *
                    (vi - vp);
          }
          da += as[0] + as[1] + as[2] + as[3];
          V3<T> f[4], tq[4];
          for (int p = 0; p < 4; ++p) {
            f[p] = (-c.mass) * as[p];
            tq[p] = cross(xs[p] - r_x, f[p]);
          }
          fsum = f[0] + f[1] + f[2] + f[3];
          tsum = tq[0] + tq[1] + tq[2] + tq[3];
          size_t o = size_t(b) * max_n + i;
          force[o] += fsum;
          torque[o] += tsum;
        }
      }
      a[i] += da;
    }
  }

  double sum_of(const std::vector<T>& arr) const {
    double s = 0;
    for (unsigned i = 0; i < n; ++i) s += static_cast<double>(arr[i]);
    return s;
  }

  T pile_cfl_v2() const {
    // Pile::calculate_cfl_v2 (pile.hpp:816-825)
    T m = std::numeric_limits<T>::lowest();
    for (const auto& B : bnd) {
      T v2 = length_sqr(cross(B.omega, V3<T>{B.sdf.max_distance, 0, 0}) + B.v);
      if (v2 > m) m = v2;
    }
    return m;
  }

  void update_dt() {
    // Solver::update_dt (solver.hpp:153-169) + calculate_cfl_v2
    T pmax = std::numeric_limits<T>::lowest();
    for (unsigned i = 0; i < n; ++i) {
      cfl_v2[i] = length_sqr(v[i] + a[i] * dt);
      pmax = std::max(pmax, cfl_v2[i]);
    }
    T pile_max = pile_cfl_v2();
    T max_v2 = std::max(pmax, pile_max);
    T max_v = std::sqrt(max_v2);
    T length_scale = g_plant == kPlantCflRadius ? c.radius : c.radius * 2;
    T cfl_dt = static_cast<T>(set.cfl) * length_scale / max_v;
    dt = std::max(std::min(cfl_dt, static_cast<T>(set.max_dt)),
                  static_cast<T>(set.min_dt));
    rep.particle_max_v2 = pmax;
    rep.pile_max_v2 = pile_max;
    rep.max_v2 = max_v2;
    rep.cfl_dt = cfl_dt;
    rep.utilized_cfl = dt * max_v / length_scale;
  }

  void update_particle_neighbors() {
    // Solver::update_particle_neighbors (solver.hpp:225-260)
    update_particle_grid();
    if (pellet_mode) {
      make_bead_pellet_lists();
      compute_density_with_pellets();
      return;
    }
    make_neighbor_list();
    compute_density();
  }

  void clear_force() {
    std::fill(force.begin(), force.end(), V3<T>{0, 0, 0});
    std::fill(torque.begin(), torque.end(), V3<T>{0, 0, 0});
  }

  void rigid_sums_and_integrate() {
    // rigid force sums (solver_df.hpp:347-353) and Pile::integrate_kinematics
    // (pile.hpp:794-815)
    for (unsigned b = 0; b < bnd.size(); ++b) {
      Boundary<T>& B = bnd[b];
      if (B.mass == 0) continue;
      double fs[3] = {0, 0, 0}, ts[3] = {0, 0, 0};
      for (unsigned i = 0; i < n; ++i) {
        size_t o = size_t(b) * max_n + i;
        fs[0] += force[o].x;
        fs[1] += force[o].y;
        fs[2] += force[o].z;
        ts[0] += torque[o].x;
        ts[1] += torque[o].y;
        ts[2] += torque[o].z;
      }
      B.force = {T(fs[0]), T(fs[1]), T(fs[2])};
      B.torque = {T(ts[0]), T(ts[1]), T(ts[2])};
    }
    for (unsigned b = 0; b < bnd.size(); ++b) integrate_kinematics(bnd[b]);
    // Pile::find_contacts + solve_contacts (solver_df.hpp:356-357)
    bool any = false;
    for (const auto& B : bnd) any |= !B.coll.empty();
    if (any) {
      if (pellet_mode)
        find_contacts_pellets();
      else
        find_contacts(-1, -1);
      solve_contacts();
    }
  }

  // ---- rigid contacts (runner.hpp:3833-3914, pile.hpp:826-979) -------------
  // R diag(v) R^T stored as diagonal + (01, 02, 12) (runner.hpp:178-199)
  static void congruent(V3<T> v, V4<T> q, V3<T>& d, V3<T>& o) {
    T m00 = 1 - 2 * (q.y * q.y + q.z * q.z), m01 = 2 * (q.x * q.y - q.z * q.w),
      m02 = 2 * (q.x * q.z + q.y * q.w);
    T m10 = 2 * (q.x * q.y + q.z * q.w), m11 = 1 - 2 * (q.x * q.x + q.z * q.z),
      m12 = 2 * (q.y * q.z - q.x * q.w);
    T m20 = 2 * (q.x * q.z - q.y * q.w), m21 = 2 * (q.y * q.z + q.x * q.w),
      m22 = 1 - 2 * (q.x * q.x + q.y * q.y);
    d = {v.x * m00 * m00 + v.y * m01 * m01 + v.z * m02 * m02,
         v.x * m10 * m10 + v.y * m11 * m11 + v.z * m12 * m12,
         v.x * m20 * m20 + v.y * m21 * m21 + v.z * m22 * m22};
    o = {v.x * m00 * m10 + v.y * m01 * m11 + v.z * m02 * m12,
         v.x * m00 * m20 + v.y * m01 * m21 + v.z * m02 * m22,
         v.x * m10 * m20 + v.y * m11 * m21 + v.z * m12 * m22};
  }
  // K = I/m - [r]x A [r]x for symmetric A = (d, o) (runner.hpp:265-288)
  static void contact_k(V3<T> r, T mass, V3<T> d, V3<T> o, V3<T>& kd, V3<T>& ko) {
    kd = {0, 0, 0};
    ko = {0, 0, 0};
    if (mass == 0) return;
    T im = 1 / mass;
    kd = {r.z * r.z * d.y - r.y * r.z * o.z * 2 + r.y * r.y * d.z + im,
          r.z * r.z * d.x - r.x * r.z * o.y * 2 + r.x * r.x * d.z + im,
          r.y * r.y * d.x - r.x * r.y * o.x * 2 + r.x * r.x * d.y + im};
    ko = {-r.z * r.z * o.x + r.x * r.z * o.z + r.y * r.z * o.y - r.x * r.y * d.z,
          r.y * r.z * o.x - r.x * r.z * d.y - r.y * r.y * o.y + r.x * r.y * o.z,
          -(r.y * r.z * d.x) + r.x * r.z * o.x + r.x * r.y * o.y - r.x * r.x * o.z};
  }
  static V3<T> sym_apply(V3<T> v, V3<T> d, V3<T> o) {
    return {v.x * d.x + v.y * o.x + v.z * o.y, v.x * o.x + v.y * d.y + v.z * o.z,
            v.x * o.y + v.y * o.z + v.z * d.z};
  }
  // collision_test (analytic distance): vertices of body i against body j
  void collision_test(unsigned i, unsigned j) {
    const Boundary<T>& Bi = bnd[i];
    const Boundary<T>& Bj = bnd[j];
    for (const V3<T>& vert : Bi.coll) {
      Contact<T> C{};
      C.cp_i = rotate(vert, Bi.q) + Bi.x;
      V3<T> local = rotate(C.cp_i - Bj.x, conjugate(Bj.q));
      V3<T> nn;
      T dist = Bj.contact_dist_normal(local, contact_tolerance, nn);
      if (!(dist < 0)) continue;
      ++num_found;
      C.i = i;
      C.j = j;
      C.cp_j = rotate(local - dist * nn, Bj.q) + Bj.x;
      C.n = rotate(nn, Bj.q);
      C.friction = Bi.friction + Bj.friction;
      V3<T> r_i = C.cp_i - Bi.x, r_j = C.cp_j - Bj.x;
      V3<T> u_rel = (Bi.v + cross(Bi.omega, r_i)) - (Bj.v + cross(Bj.omega, r_j));
      T u_rel_n = dot(C.n, u_rel);
      C.t = u_rel - u_rel_n * C.n;
      if (length_sqr(C.t) > static_cast<T>(1e-6)) C.t = C.t * (1 / std::sqrt(length_sqr(C.t)));
      congruent({1 / Bi.inertia.x, 1 / Bi.inertia.y, 1 / Bi.inertia.z}, Bi.q, C.iiwi_diag,
                C.iiwi_off_diag);
      congruent({1 / Bj.inertia.x, 1 / Bj.inertia.y, 1 / Bj.inertia.z}, Bj.q, C.iiwj_diag,
                C.iiwj_off_diag);
      V3<T> kdi, koi, kdj, koj;
      contact_k(r_i, Bi.mass, C.iiwi_diag, C.iiwi_off_diag, kdi, koi);
      contact_k(r_j, Bj.mass, C.iiwj_diag, C.iiwj_off_diag, kdj, koj);
      V3<T> kd = kdi + kdj, ko = koi + koj;
      C.nkninv = 1 / dot(C.n, sym_apply(C.n, kd, ko));
      C.pmax = 1 / dot(C.t, sym_apply(C.t, kd, ko)) * dot(u_rel, C.t);
      C.goalu = u_rel_n < 0 ? -(Bi.restitution * Bj.restitution) * u_rel_n : T(0);
      C.impulse_sum = 0;
      if (contacts.size() < max_num_contacts) contacts.push_back(C);
    }
  }
  // i < 0: all pairs; j < 0: body i against all others (skipping i == j and
  // pairs of two massless bodies); box-shell bodies j have no collision test
  void find_contacts(int i, int j) {
    num_found = 0;
    contacts.clear();
    unsigned nb = unsigned(bnd.size());
    unsigned i0 = i < 0 ? 0 : unsigned(i), i1 = i < 0 ? nb : unsigned(i) + 1;
    for (unsigned a = i0; a < i1; ++a) {
      unsigned j0 = j < 0 ? 0 : unsigned(j), j1 = j < 0 ? nb : unsigned(j) + 1;
      for (unsigned b = j0; b < j1; ++b) {
        if (j < 0 && (a == b || (bnd[a].mass == 0 && bnd[b].mass == 0))) continue;
        if (bnd[b].sdf.type == ORC_BOX_SHELL) continue;
        collision_test(a, b);
      }
    }
  }
  // solve_contacts (pile.hpp:906-979)
  void solve_contacts() {
    for (int sweep = 0; sweep < 5; ++sweep) {
      for (Contact<T>& C : contacts) {
        Boundary<T>& Bi = bnd[C.i];
        Boundary<T>& Bj = bnd[C.j];
        T depth = dot(C.n, C.cp_i - C.cp_j);
        V3<T> r_i = C.cp_i - Bi.x, r_j = C.cp_j - Bj.x;
        V3<T> u_rel = (Bi.v + cross(Bi.omega, r_i)) - (Bj.v + cross(Bj.omega, r_j));
        T u_rel_n = dot(C.n, u_rel);
        T corr = C.nkninv * (C.goalu - u_rel_n);
        if (corr < -C.impulse_sum) corr = -C.impulse_sum;
        if (depth < 0) corr -= C.nkninv * depth;
        V3<T> p = corr * C.n;
        C.impulse_sum += corr;
        T pn = dot(p, C.n);
        if (C.friction * pn > C.pmax)
          p = p - C.pmax * C.t;
        else if (C.friction * pn < -C.pmax)
          p = p + C.pmax * C.t;
        else
          p = p - (C.friction * pn) * C.t;
        if (Bi.mass != 0) {
          Bi.v += p * (1 / Bi.mass);
          Bi.omega += sym_apply(cross(r_i, p), C.iiwi_diag, C.iiwi_off_diag);
        }
        if (Bj.mass != 0) {
          Bj.v += (-p) * (1 / Bj.mass);
          Bj.omega += sym_apply(cross(r_j, -p), C.iiwj_diag, C.iiwj_off_diag);
        }
      }
    }
  }

  static V4<T> dq_of(V3<T> w, V4<T> q) {
    // calculate_dq (runner.hpp:241-252)
    return {static_cast<T>(0.5) * (w.x * q.w + w.y * q.z - w.z * q.y),
            static_cast<T>(0.5) * (-w.x * q.z + w.y * q.w + w.z * q.x),
            static_cast<T>(0.5) * (w.x * q.y - w.y * q.x + w.z * q.w),
            static_cast<T>(0.5) * (-w.x * q.x - w.y * q.y - w.z * q.z)};
  }
  static V4<T> qnormalize(V4<T> q) {
    T s = 1 / std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return {q.x * s, q.y * s, q.z * s, q.w * s};
  }
  static V3<T> angular_acceleration(V3<T> inertia, V4<T> q, V3<T> tau) {
    // calculate_angular_acceleration (runner.hpp:202-239): R diag(1/I) R^T tau
    V3<T> local = rotate(tau, conjugate(q));
    V3<T> scaled{local.x / inertia.x, local.y / inertia.y, local.z / inertia.z};
    return rotate(scaled, q);
  }
  void integrate_kinematics(Boundary<T>& B) {
    if (B.mass == 0) {
      B.x += B.v * dt;
      V4<T> dq = dq_of(B.omega, B.q);
      B.q = {B.q.x + dt * dq.x, B.q.y + dt * dq.y, B.q.z + dt * dq.z,
             B.q.w + dt * dq.w};
      B.q = qnormalize(B.q);
    } else {
      B.v += ((1 / B.mass) * B.force) * dt;
      B.omega += angular_acceleration(B.inertia, B.q, B.torque) * dt;
      B.a = pile_gravity;
      V4<T> dq = dq_of(B.omega, B.q);
      B.q = {B.q.x + dt * dq.x, B.q.y + dt * dq.y, B.q.z + dt * dq.z,
             B.q.w + dt * dq.w};
      B.q = qnormalize(B.q);
      V3<T> dx = (B.a * dt + B.v) * dt;
      B.x += dx;
      B.v = (1 / dt) * dx;
    }
  }

  // ---- DFSPH step (solver_df.hpp:74-359) ------------------------------------
  void step_df() {
    clear_force();
    compute_all_boundaries();
    update_particle_neighbors();
    if (set.enable_divergence_solve) {
      compute_dfsph_factor();
      // warm_start_divergence_solve_0 (runner.hpp:3230-3250)
      std::vector<T> kv_new(n);
      const T div_clamp = static_cast<T>(g_plant == kPlantDivWarmClamp ? -0.25 : -0.5);
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        T dadv = density_change(i);
        kv_new[i] = dadv > 0 ? static_cast<T>(0.5) * std::max(kappa_v[i], div_clamp) / dt
                             : static_cast<T>(0);
      }
      for (unsigned i = 0; i < n; ++i) kappa_v[i] = kv_new[i];
      // warm_start_divergence_solve_1 (runner.hpp:3252-3295)
      std::vector<V3<T>> dvs(n);
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        dvs[i] = kappa_dv(i, kappa_v[i], dt, true,
                          [&](unsigned j) { return kappa_v[j]; });
      }
      for (unsigned i = 0; i < n; ++i) v[i] += dvs[i];
      // compute_velocity_of_density_change (runner.hpp:3297-3314)
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        density_adv[i] = density_change(i);
        T inv_dt = 1 / dt;
        dfsph_factor[i] *= inv_dt;
      }
      T mean = std::numeric_limits<T>::max();
      unsigned it = 0;
      while ((mean > static_cast<T>(set.density_change_tolerance) / dt ||
              it < set.min_divergence_solve) &&
             it < set.max_divergence_solve) {
        // divergence_solve_iteration (runner.hpp:3316-3363)
#pragma omp parallel for schedule(static)
        for (long ii = 0; ii < long(n); ++ii) {
          unsigned i = unsigned(ii);
          T b_i = density_adv[i];
          T k_i = b_i * dfsph_factor[i];
          kappa_v[i] += k_i;
          dvs[i] = kappa_dv(i, k_i, dt, true, [&](unsigned j) {
            return density_adv[j] * dfsph_factor[j];
          });
        }
        for (unsigned i = 0; i < n; ++i) v[i] += dvs[i];
        // compute_divergence_solve_density_error (runner.hpp:3396-3410)
#pragma omp parallel for schedule(static)
        for (long ii = 0; ii < long(n); ++ii)
          density_adv[ii] = density_change(unsigned(ii));
        mean = static_cast<T>(sum_of(density_adv) / double(n));
        ++it;
      }
      rep.num_divergence_solve = it;
      rep.mean_density_change = mean;
      // divergence_solve_finish (runner.hpp:3412-3420)
      for (unsigned i = 0; i < n; ++i) {
        dfsph_factor[i] *= dt;
        kappa_v[i] *= g_plant == kPlantDivFinish ? dt * dt : dt;
      }
    }
    // clear_acceleration (runner.hpp:1974-1980)
    for (unsigned i = 0; i < n; ++i) a[i] = c.g;
    compute_viscosity();
    update_dt();
    // integrate_non_pressure_acceleration (runner.hpp:3422-3428)
    for (unsigned i = 0; i < n; ++i) v[i] += dt * a[i];
    if (set.enable_density_solve) {
      // warm_start_pressure_solve0 (runner.hpp:3430-3450)
      std::vector<T> k_new(n);
      const T dens_clamp = static_cast<T>(g_plant == kPlantDensWarmClamp ? -0.0005 : -0.00025);
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        T dadv = density_adv_of(i, dt);
        k_new[i] = dadv > 1.0 ? (static_cast<T>(0.5) * std::max(kappa[i], dens_clamp) / (dt * dt))
                              : static_cast<T>(0);
      }
      for (unsigned i = 0; i < n; ++i) kappa[i] = k_new[i];
      // warm_start_pressure_solve1 (runner.hpp:3452-3491)
      std::vector<V3<T>> dvs(n);
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        dvs[i] = kappa_dv(i, kappa[i], dt, false,
                          [&](unsigned j) { return kappa[j]; });
      }
      for (unsigned i = 0; i < n; ++i) v[i] += dvs[i];
      // compute_rho_adv (runner.hpp:3525-3541)
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        density_adv[i] = density_adv_of(i, dt);
        dfsph_factor[i] *= 1 / (dt * dt);
      }
      T mean = std::numeric_limits<T>::max();
      unsigned it = 0;
      while ((mean > static_cast<T>(set.density_error_tolerance) ||
              it < set.min_density_solve) &&
             it < set.max_density_solve) {
        // pressure_solve_iteration (runner.hpp:3543-3588)
#pragma omp parallel for schedule(static)
        for (long ii = 0; ii < long(n); ++ii) {
          unsigned i = unsigned(ii);
          T b_i = density_adv[i] - 1;
          T k_i = b_i * dfsph_factor[i];
          kappa[i] += k_i;
          dvs[i] = kappa_dv(i, k_i, dt, false, [&](unsigned j) {
            return (density_adv[j] - 1) * dfsph_factor[j];
          });
        }
        for (unsigned i = 0; i < n; ++i) v[i] += dvs[i];
        // compute_pressure_solve_density_error (runner.hpp:3590-3604)
#pragma omp parallel for schedule(static)
        for (long ii = 0; ii < long(n); ++ii)
          density_adv[ii] = density_adv_of(unsigned(ii), dt);
        double s = 0;
        for (unsigned i = 0; i < n; ++i) s += double(density_adv[i] - 1);
        mean = static_cast<T>(s / double(n));
        ++it;
      }
      rep.num_density_solve = it;
      rep.mean_density_error = mean;
      // pressure_solve_finish (runner.hpp:3606-3611)
      for (unsigned i = 0; i < n; ++i) kappa[i] *= g_plant == kPlantDensFinish ? dt : dt * dt;
    }
    integrate_velocity();
    rigid_sums_and_integrate();
    t += dt;
  }

  void integrate_velocity() {
    // integrate_velocity<wrap> (runner.hpp:3613-3623), wrap per enabled axis
    for (unsigned i = 0; i < n; ++i) x[i] = wrap_pos(x[i] + dt * v[i]);
  }
  V3<T> wrap_pos(V3<T> p) const { return wrap_disp(p); }

  // ---- IISPH step (solver_ii.hpp:76-243) -------------------------------------
  void step_ii() {
    clear_force();
    for (unsigned i = 0; i < n; ++i) a[i] = c.g;
    compute_all_boundaries();
    update_particle_neighbors();
    compute_viscosity();
    update_dt();
    // predict_advection0 (runner.hpp:2609-2639)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      v[i] += dt * a[i];
      last_pressure[i] = pressure[i] * static_cast<T>(0.5);
      T dens = density[i];
      T dens2 = g_plant == kPlantIiDiiDensity ? dens * c.rho0 : dens * dens;
      V3<T> d{0, 0, 0};
      for (unsigned s = 0; s < num_neighbors[i]; ++s)
        d -= (c.vol / dens2) *
             displacement_cubic_kernel_grad(nb_x[size_t(i) * c.kmax + s], c);
      for (unsigned b = 0; b < bnd.size(); ++b) {
        V4<T> k = bker(b, i);
        V3<T> g{k.x, k.y, k.z};
        d -= V3<T>{g.x / dens2, g.y / dens2, g.z / dens2};
      }
      dii[i] = d * c.rho0;
    }
    // predict_advection1 (runner.hpp:2640-2690)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      V3<T> vi = v[i], d_ii = dii[i];
      T dens = density[i];
      T dens2 = dens * dens;
      T inv_dens2 = 1 / dens2;
      T dpi = c.vol * inv_dens2;
      T dadv = dens / c.rho0;
      T a_ii = 0;
      for (unsigned s = 0; s < num_neighbors[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = nb_id[o];
        V3<T> grad_w = displacement_cubic_kernel_grad(nb_x[o], c);
        V3<T> dji = (dpi * c.rho0) * grad_w;
        dadv += dt * c.vol * dot(vi - v[j], grad_w);
        a_ii += c.vol * dot(g_plant == kPlantIiAiiNoDji ? d_ii : d_ii - dji, grad_w);
      }
      for (unsigned b = 0; b < bnd.size(); ++b) {
        V4<T> r = brec(b, i), k = bker(b, i);
        V3<T> g{k.x, k.y, k.z};
        V3<T> velj = rigid_vel(b, {r.x, r.y, r.z});
        V3<T> dji = (inv_dens2 * c.rho0) * g;
        dadv += dt * dot(vi - velj, g);
        a_ii += dot(d_ii - dji, g);
      }
      adv_density[i] = dadv;
      aii[i] = a_ii * c.rho0;
    }
    T mean = std::numeric_limits<T>::max();
    unsigned it = 0;
    while ((mean > static_cast<T>(set.density_error_tolerance) ||
            it < set.min_density_solve) &&
           it < set.max_density_solve) {
      // pressure_solve_iteration0 (runner.hpp:2691-2715)
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        V3<T> s3{0, 0, 0};
        for (unsigned s = 0; s < num_neighbors[i]; ++s) {
          size_t o = size_t(i) * c.kmax + s;
          unsigned j = nb_id[o];
          T dj = density[j];
          T dj2 = dj * dj;
          s3 -= (c.vol / dj2 * last_pressure[j]) *
                displacement_cubic_kernel_grad(nb_x[o], c);
        }
        dij_pj[i] = s3 * c.rho0;
      }
      // pressure_solve_iteration1 (runner.hpp:2716-2772)
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        T lp = last_pressure[i];
        V3<T> dpj = dij_pj[i];
        T dens = density[i];
        T dens2 = dens * dens;
        T dpi = c.vol / dens2;
        T bb = 1 - adv_density[i];
        T dt2 = dt * dt;
        T a_ii = aii[i];
        T denom = a_ii * dt2;
        const T w = g_plant == kPlantIiOmega ? static_cast<T>(0.5) : static_cast<T>(2) / 3;
        T p = 0;
        T sum_tmp = 0;
        for (unsigned s = 0; s < num_neighbors[i]; ++s) {
          size_t o = size_t(i) * c.kmax + s;
          unsigned j = nb_id[o];
          V3<T> djk_pk = dij_pj[j];
          V3<T> grad_w = displacement_cubic_kernel_grad(nb_x[o], c);
          V3<T> dji = (dpi * c.rho0) * grad_w;
          V3<T> dji_pi = dji * lp;
          sum_tmp += c.vol * dot(dpj - dii[j] * last_pressure[j] - (djk_pk - dji_pi),
                                 grad_w);
        }
        for (unsigned b = 0; b < bnd.size(); ++b) {
          V4<T> k = bker(b, i);
          sum_tmp += dot(dpj, V3<T>{k.x, k.y, k.z});
        }
        sum_tmp *= c.rho0;
        if (std::fabs(denom) > static_cast<T>(1.0e-9))
          p = std::max((1 - w) * lp + w / denom * (bb - dt2 * sum_tmp),
                       static_cast<T>(0));
        density_err[i] = p == 0 ? 0 : (denom * p + sum_tmp * dt2 - bb);
        pressure[i] = p;
      }
      for (unsigned i = 0; i < n; ++i) last_pressure[i] = pressure[i];
      mean = static_cast<T>(sum_of(density_err) / double(n));
      ++it;
    }
    rep.num_density_solve = it;
    rep.mean_density_error = mean;
    pressure_accels_and_integrate();
  }

  // compute_pressure_accels, kinematic_integration, rigid sums: the end of
  // SolverIi::step and SolverI::step (solver_ii.hpp:208-241,
  // solver_i.hpp:289-327)
  void pressure_accels_and_integrate() {
    compute_pressure_accels();
    // kinematic_integration<wrap> (runner.hpp:2873-2887)
    for (unsigned i = 0; i < n; ++i) {
      V3<T> vv = v[i];
      vv += pressure_accel[i] * dt;
      x[i] = wrap_pos(x[i] + vv * dt);
      v[i] = vv;
    }
    rigid_sums_and_integrate();
    t += dt;
  }

  void compute_pressure_accels() {
    // compute_pressure_accels (runner.hpp:2774-2817)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      T dens = density[i];
      T dens2 = dens * dens;
      T dpi = pressure[i] / dens2;
      V3<T> ai{0, 0, 0};
      for (unsigned s = 0; s < num_neighbors[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = nb_id[o];
        T dj = density[j];
        T dj2 = dj * dj;
        T dpj = pressure[j] / dj2;
        ai -= (c.mass * (dpi + dpj)) * displacement_cubic_kernel_grad(nb_x[o], c);
      }
      if (pellet_mode) {
        // compute_pressure_accels_with_pellets (runner.hpp:2819-2870)
        for (unsigned s = 0; s < num_pnb[i]; ++s) {
          size_t o = size_t(i) * c.kmax + s;
          unsigned j = pnb_id[o];
          V3<T> xixj = pnb_x[o];
          V3<T> xj = x[i] - xixj;
          unsigned b = pellet_body[j - max_n];
          T dj2 = c.rho0 * c.rho0;
          T dpj = pressure[j] / dj2;
          V3<T> acc = (c.mass * (dpi + dpj)) * displacement_cubic_kernel_grad(xixj, c);
          V3<T> f = c.mass * acc;
          ai -= acc;
          size_t ob = size_t(b) * max_n + i;
          force[ob] += f;
          torque[ob] += cross(xj - bnd[b].x, f);
        }
      } else {
        for (unsigned b = 0; b < bnd.size(); ++b) {
          V4<T> r = brec(b, i), k = bker(b, i);
          V3<T> acc = (c.rho0 * dpi) * V3<T>{k.x, k.y, k.z};
          V3<T> f = c.mass * acc;
          ai -= acc;
          add_force(b, i, {r.x, r.y, r.z}, f);
        }
      }
      pressure_accel[i] = c.rho0 * ai;
    }
  }

  // ---- SolverI::step (solver_i.hpp:79-330), volume-map boundaries ------------
  void step_i() {
    if (pellet_mode) {
      step_i_pellets();
      return;
    }
    clear_force();
    for (unsigned i = 0; i < n; ++i) a[i] = c.g;
    compute_all_boundaries();
    update_particle_neighbors();
    compute_viscosity();
    update_dt();
    // advect_and_init_pressure (runner.hpp:2890-2900)
    for (unsigned i = 0; i < n; ++i) {
      v[i] += dt * a[i];
      last_pressure[i] = pressure[i] * static_cast<T>(0.5);
    }
    const T eps = static_cast<T>(0.01) * c.h2;
    // calculate_isph_diagonal_adv_density (runner.hpp:2901-2949); the
    // diagonal times rho0 is kept in aii
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      V3<T> xi = x[i], vi = v[i];
      T dens = density[i];
      T dadv = dens / c.rho0;
      T diag = 0;
      for (unsigned s = 0; s < num_neighbors[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = nb_id[o];
        V3<T> xixj = nb_x[o];
        V3<T> grad_w = displacement_cubic_kernel_grad(xixj, c);
        T dj = density[j];
        dadv += dt * c.vol * dot(vi - v[j], grad_w);
        diag += c.vol * (dens + dj) / (dens * dj) / (length_sqr(xixj) + eps) * dot(xixj, grad_w);
      }
      for (unsigned b = 0; b < bnd.size(); ++b) {
        V4<T> r = brec(b, i), k = bker(b, i);
        V3<T> g{k.x, k.y, k.z};
        V3<T> bx{r.x, r.y, r.z};
        V3<T> xib = xi - bx;
        dadv += dt * dot(vi - rigid_vel(b, bx), g);
        diag += (dens + c.rho0) / (dens * c.rho0) / (length_sqr(xib) + eps) * dot(xib, g);
      }
      aii[i] = diag * c.rho0;
      adv_density[i] = dadv;
    }
    T mean = std::numeric_limits<T>::max();
    unsigned it = 0;
    while ((mean > static_cast<T>(set.density_error_tolerance) ||
            it < set.min_density_solve) &&
           it < set.max_density_solve) {
      // isph_solve_iteration (runner.hpp:2996-3042), weighted Jacobi
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        T dens = density[i];
        T diag = aii[i];
        T bb = 1 - adv_density[i];
        T dt2 = dt * dt;
        T denom = diag * dt2;
        const T w = static_cast<T>(2) / 3;
        T off = 0;
        for (unsigned s = 0; s < num_neighbors[i]; ++s) {
          size_t o = size_t(i) * c.kmax + s;
          unsigned j = nb_id[o];
          V3<T> xixj = nb_x[o];
          T dj = density[j];
          off -= c.vol * (dens + dj) / (dens * dj) / (length_sqr(xixj) + eps) *
                 dot(xixj, displacement_cubic_kernel_grad(xixj, c)) * last_pressure[j];
        }
        off *= c.rho0;
        T p = 0;
        if (std::fabs(denom) > static_cast<T>(1.0e-9))
          p = std::max((1 - w) * last_pressure[i] + w * (bb / denom - off / diag),
                       static_cast<T>(0));
        density_err[i] = p == 0 ? 0 : (denom * p + off * dt2 - bb);
        pressure[i] = p;
      }
      for (unsigned i = 0; i < n; ++i) last_pressure[i] = pressure[i];
      mean = static_cast<T>(sum_of(density_err) / double(n));
      ++it;
    }
    rep.num_density_solve = it;
    rep.mean_density_error = mean;
    pressure_accels_and_integrate();
  }

  // ---- particle-sampled boundaries (pellets) ----------------------------------
  // pellet_id_to_rigid_id from the bodies' pellet counts (pile.hpp:624-636)
  int refresh_pellets() {
    pellet_body.clear();
    for (unsigned b = 0; b < bnd.size(); ++b)
      for (size_t k = 0; k < bnd[b].coll.size(); ++k) pellet_body.push_back(b);
    if (pellet_body.size() > max_p) return -4;  // pile.hpp:266-275
    np = unsigned(pellet_body.size());
    return 0;
  }
  void snapshot_kinematics() {
    kin0.resize(bnd.size());
    for (unsigned b = 0; b < bnd.size(); ++b) kin0[b] = {bnd[b].x, bnd[b].v, bnd[b].omega, bnd[b].q};
  }
  // transform_all_pellets (solver.hpp:204-223, transform_pellets
  // runner.hpp:4071-4085) with the step-start kinematics
  void transform_all_pellets() {
    unsigned k = 0;
    for (unsigned b = 0; b < bnd.size(); ++b) {
      const Kin& K0 = kin0[b];
      for (const V3<T>& local : bnd[b].coll) {
        V3<T> r = rotate(local, K0.q);
        x[max_n + k] = r + K0.x;
        v[max_n + k] = cross(K0.omega, r) + K0.v;
        ++k;
      }
    }
  }
  // make_bead_pellet_neighbor_list<wrap> (runner.hpp:1913-1971) over the
  // fluid particles and the pellets (..._check_contiguous,
  // runner.hpp:5642-5666): the 27 cells in the reference order, fluid
  // neighbours into the bead list and pellets into the pellet list
  void make_bead_pellet_lists() {
    const unsigned K = c.kmax;
    int flag2 = 0;
    const long total = long(n) + long(np);
#pragma omp parallel for schedule(static) reduction(| : flag2)
    for (long kk = 0; kk < total; ++kk) {
      const unsigned i = kk < long(n) ? unsigned(kk) : max_n + unsigned(kk - long(n));
      V3<T> xi = x[i];
      int ip[3];
      ipos_of(xi, ip);
      unsigned cb = 0, cp = 0;
      for (int t = 0; t < 27; ++t) {
        int nq[3] = {ip[0] + t / 9 - 1, ip[1] + (t / 3) % 3 - 1, ip[2] + t % 3 - 1};
        for (int ax = 0; ax < 3; ++ax) {
          if (!(c.wrap_mask & (1 << ax))) continue;
          if (nq[ax] >= c.res[ax])
            nq[ax] -= c.res[ax];
          else if (nq[ax] < 0)
            nq[ax] += c.res[ax];
        }
        if (!within_grid(nq)) continue;
        unsigned cell = linear(nq);
        for (unsigned q = cell_start[cell]; q < cell_start[cell + 1]; ++q) {
          unsigned j = cell_ids[q];
          V3<T> xixj = wrap_disp(xi - x[j]);
          if (j == i || !(length_sqr(xixj) < c.h2)) continue;
          if (j < max_n) {
            size_t slot = size_t(i) * K + std::min(K - 1, cb);
            nb_id[slot] = j;
            nb_x[slot] = xixj;
            ++cb;
          } else {
            size_t slot = size_t(i) * K + std::min(K - 1, cp);
            pnb_id[slot] = j;
            pnb_x[slot] = xixj;
            ++cp;
          }
        }
      }
      num_neighbors[i] = std::min(K, cb);
      num_pnb[i] = std::min(K, cp);
      if (cb > K || cp > K) flag2 = 1;
    }
    if (flag2) anomaly[2] = 1;
  }
  // compute_density_with_pellets (runner.hpp:2074-2098)
  void compute_density_with_pellets() {
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      T dens = c.w0;
      for (unsigned s = 0; s < num_neighbors[i]; ++s)
        dens += displacement_cubic_kernel(nb_x[size_t(i) * c.kmax + s], c);
      for (unsigned s = 0; s < num_pnb[i]; ++s)
        dens += displacement_cubic_kernel(pnb_x[size_t(i) * c.kmax + s], c);
      density[i] = dens * c.mass;
    }
  }
  // compute_viscosity_with_pellets (runner.hpp:2251-2297)
  void compute_viscosity_with_pellets() {
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      V3<T> vi = v[i], xi = x[i];
      V3<T> da{0, 0, 0};
      for (unsigned s = 0; s < num_neighbors[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = nb_id[o];
        V3<T> xixj = nb_x[o];
        da += (c.visc * c.mass / density[j] * dot(xixj, displacement_cubic_kernel_grad(xixj, c)) /
               (length_sqr(xixj) + static_cast<T>(0.01) * c.h2)) *
              (vi - v[j]);
      }
      for (unsigned s = 0; s < num_pnb[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = pnb_id[o];
        V3<T> xixj = pnb_x[o];
        V3<T> acc = (c.vol * c.bvisc * dot(xixj, displacement_cubic_kernel_grad(xixj, c)) /
                     (length_sqr(xixj) + static_cast<T>(0.01) * c.h2)) *
                    (vi - v[j]);
        da += acc;
        V3<T> px = xi - xixj;
        unsigned b = pellet_body[j - max_n];
        V3<T> f = (-c.mass) * acc;
        size_t ob = size_t(b) * max_n + i;
        force[ob] += f;
        torque[ob] += cross(px - bnd[b].x, f);
      }
      a[i] += da;
    }
  }
  // Pseudo-inverse of the symmetric 3x3 moment matrix of the MLS plane fit
  // (runner.hpp:2581-2590 takes it from an SVD, runner.hpp:340-574).  For a
  // symmetric matrix V S^+ U^T equals sum_k v_k v_k^T / lambda_k over the
  // eigenpairs with |lambda_k| > 1e-2, the reference's singular-value cut.
  // Eigenpairs by cyclic Jacobi rotations (Jacobi 1846; the rotation
  // formulas of Golub and Van Loan, Matrix Computations, 8.5.2), six sweeps.
  static void sym_pinv(V3<T> d, V3<T> o, T inv[3][3]) {
    T m[3][3] = {{d.x, o.x, o.y}, {o.x, d.y, o.z}, {o.y, o.z, d.z}};
    T e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 6; ++sweep) {
      for (int r = 0; r < 3; ++r) {
        const int pp = P[r], qq = Q[r];
        const T apq = m[pp][qq];
        if (apq == 0) continue;
        const T theta = (m[qq][qq] - m[pp][pp]) / (2 * apq);
        const T tn = (theta >= 0 ? T(1) : T(-1)) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const T cs = 1 / std::sqrt(tn * tn + 1), sn = tn * cs;
        m[pp][pp] -= tn * apq;
        m[qq][qq] += tn * apq;
        m[pp][qq] = m[qq][pp] = 0;
        const int k = 3 - pp - qq;
        const T akp = m[k][pp], akq = m[k][qq];
        m[k][pp] = m[pp][k] = cs * akp - sn * akq;
        m[k][qq] = m[qq][k] = sn * akp + cs * akq;
        for (int w = 0; w < 3; ++w) {
          const T ep = e[w][pp], eq = e[w][qq];
          e[w][pp] = cs * ep - sn * eq;
          e[w][qq] = sn * ep + cs * eq;
        }
      }
    }
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) inv[r][cc] = 0;
    for (int k = 0; k < 3; ++k) {
      const T lam = m[k][k];
      if (!(std::fabs(lam) > static_cast<T>(1e-2))) continue;
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) inv[r][cc] += e[r][k] * e[cc][k] / lam;
    }
  }
  // extrapolate_pressure_to_pellet (runner.hpp:2543-2607): weighted plane
  // fit of the fluid neighbours' pressure, clamped at 0 (fmax: a pellet
  // without fluid neighbours gets 0)
  void extrapolate_pressure(std::vector<T>& pr) {
#pragma omp parallel for schedule(static)
    for (long kk = 0; kk < long(np); ++kk) {
      const unsigned i = max_n + unsigned(kk);
      const unsigned nn = num_neighbors[i];
      T w_sum = 0;
      V3<T> center{0, 0, 0};
      for (unsigned s = 0; s < nn; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        T w = displacement_cubic_kernel(nb_x[o], c) * c.vol;
        w_sum += w;
        center += w * x[nb_id[o]];
      }
      center = V3<T>{center.x / w_sum, center.y / w_sum, center.z / w_sum};
      T alpha = 0;
      V3<T> dg{0, 0, 0}, od{0, 0, 0}, pxs{0, 0, 0};
      for (unsigned s = 0; s < nn; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = nb_id[o];
        V3<T> xt = x[j] - center;
        T w = displacement_cubic_kernel(nb_x[o], c) * c.vol;
        T wp = w * pr[j];
        alpha += wp;
        pxs += wp * xt;
        dg += w * V3<T>{xt.x * xt.x, xt.y * xt.y, xt.z * xt.z};
        od += w * V3<T>{xt.x * xt.y, xt.x * xt.z, xt.y * xt.z};
      }
      alpha /= w_sum;
      T inv[3][3];
      sym_pinv(dg, od, inv);
      V3<T> plane{inv[0][0] * pxs.x + inv[0][1] * pxs.y + inv[0][2] * pxs.z,
                  inv[1][0] * pxs.x + inv[1][1] * pxs.y + inv[1][2] * pxs.z,
                  inv[2][0] * pxs.x + inv[2][1] * pxs.y + inv[2][2] * pxs.z};
      T p = alpha + dot(plane, x[i] - center);
      pr[i] = std::fmax(p, static_cast<T>(0));
    }
  }
  // collision_test_with_pellets (runner.hpp:3627-3727): a world pellet of
  // body i against body j, with the device kinematics of the step start
  void collision_test_pellet(unsigned k, unsigned j) {
    const unsigned i = pellet_body[k];
    if (i == j || (bnd[i].mass == 0 && bnd[j].mass == 0)) return;
    const Boundary<T>& Bi = bnd[i];
    const Boundary<T>& Bj = bnd[j];
    const Kin &Ki = kin0[i], &Kj = kin0[j];
    Contact<T> C{};
    C.cp_i = x[max_n + k];
    V3<T> local = rotate(C.cp_i - Kj.x, conjugate(Kj.q));
    V3<T> nn;
    T dist = Bj.contact_dist_normal(local, contact_tolerance, nn);
    if (!(dist < 0)) return;
    ++num_found;
    C.i = i;
    C.j = j;
    C.cp_j = rotate(local - dist * nn, Kj.q) + Kj.x;
    C.n = rotate(nn, Kj.q);
    C.friction = Bi.friction + Bj.friction;
    V3<T> r_i = C.cp_i - Ki.x, r_j = C.cp_j - Kj.x;
    V3<T> u_rel = (Ki.v + cross(Ki.omega, r_i)) - (Kj.v + cross(Kj.omega, r_j));
    T u_rel_n = dot(C.n, u_rel);
    C.t = u_rel - u_rel_n * C.n;
    if (length_sqr(C.t) > static_cast<T>(1e-6)) C.t = C.t * (1 / std::sqrt(length_sqr(C.t)));
    congruent({1 / Bi.inertia.x, 1 / Bi.inertia.y, 1 / Bi.inertia.z}, Ki.q, C.iiwi_diag,
              C.iiwi_off_diag);
    congruent({1 / Bj.inertia.x, 1 / Bj.inertia.y, 1 / Bj.inertia.z}, Kj.q, C.iiwj_diag,
              C.iiwj_off_diag);
    V3<T> kdi, koi, kdj, koj;
    contact_k(r_i, Bi.mass, C.iiwi_diag, C.iiwi_off_diag, kdi, koi);
    contact_k(r_j, Bj.mass, C.iiwj_diag, C.iiwj_off_diag, kdj, koj);
    V3<T> kd = kdi + kdj, ko = koi + koj;
    C.nkninv = 1 / dot(C.n, sym_apply(C.n, kd, ko));
    C.pmax = 1 / dot(C.t, sym_apply(C.t, kd, ko)) * dot(u_rel, C.t);
    C.goalu = u_rel_n < 0 ? -(Bi.restitution * Bj.restitution) * u_rel_n : T(0);
    C.impulse_sum = 0;
    if (contacts.size() < max_num_contacts) contacts.push_back(C);
  }
  // Pile::find_contacts(particle_x, max_num_beads), pellet branch
  // (pile.hpp:858-886): one test per body j, or per hinted run of identical
  // bodies; contacts in (run, pellet, j) order; box-shell bodies have no
  // pellet test (runner.hpp:5980-6060)
  void find_contacts_pellets() {
    num_found = 0;
    contacts.clear();
    const unsigned nb = unsigned(bnd.size());
    unsigned j = 0, seq = 0;
    while (j < nb) {
      unsigned jb = j, je = j + 1;
      if (seq < ident.size() && j == ident[seq].first && j < ident[seq].second) {
        jb = ident[seq].first;
        je = std::min(nb, ident[seq].second);
        ++seq;
      }
      if (bnd[jb].sdf.type != ORC_BOX_SHELL)
        for (unsigned k = 0; k < np; ++k)
          for (unsigned jj = jb; jj < je; ++jj) collision_test_pellet(k, jj);
      j = je;
    }
  }
  // make_bead_pellet_neighbor_list<wrap> (runner.hpp:1913-1971) on sample
  // points offset .. offset+ns of sx (rows indexed by the sample's index)
  void sample_bead_pellet(const V3<T>* sx, unsigned ns, unsigned offset, unsigned max_beads,
                          unsigned* bid, V3<T>* bx, unsigned* bc, unsigned* pid, V3<T>* px,
                          unsigned* pc) const {
    const unsigned K = c.kmax;
    for (unsigned i0 = 0; i0 < ns; ++i0) {
      const unsigned i = i0 + offset;
      V3<T> xi = sx[i];
      int ip[3];
      ipos_of(xi, ip);
      unsigned cb = 0, cp = 0;
      for (int t = 0; t < 27; ++t) {
        int nq[3] = {ip[0] + t / 9 - 1, ip[1] + (t / 3) % 3 - 1, ip[2] + t % 3 - 1};
        for (int ax = 0; ax < 3; ++ax) {
          if (!(c.wrap_mask & (1 << ax))) continue;
          if (nq[ax] >= c.res[ax])
            nq[ax] -= c.res[ax];
          else if (nq[ax] < 0)
            nq[ax] += c.res[ax];
        }
        if (!within_grid(nq)) continue;
        unsigned cell = linear(nq);
        for (unsigned q = cell_start[cell]; q < cell_start[cell + 1]; ++q) {
          unsigned j = cell_ids[q];
          V3<T> d = wrap_disp(xi - grid_x[j]);
          if (j == i || !(length_sqr(d) < c.h2)) continue;
          if (j < max_beads) {
            size_t o = size_t(i) * K + std::min(K - 1, cb++);
            bid[o] = j;
            bx[o] = d;
          } else {
            size_t o = size_t(i) * K + std::min(K - 1, cp++);
            pid[o] = j;
            px[o] = d;
          }
        }
      }
      bc[i] = std::min(K, cb);
      pc[i] = std::min(K, cp);
    }
  }
  // sums over lists: 0 compute_particle_boundary_with_pellets
  // (runner.hpp:2027-2050), 1 compute_density_with_pellets
  // (runner.hpp:2074-2098), 2 compute_dfsph_factor_with_pellets
  // (runner.hpp:3156-3190, the pellet entries read as grad_wvol)
  void pellet_sums(int what, unsigned K, unsigned nrows, const V3<T>* bx, const unsigned* bc,
                   const V3<T>* px, const unsigned* pc, T* out) const {
    for (unsigned i = 0; i < nrows; ++i) {
      const unsigned nb = bc ? std::min(bc[i], K) : 0u, np_ = std::min(pc[i], K);
      if (what == 0) {
        V3<T> g{0, 0, 0};
        T w = 0;
        for (unsigned t = 0; t < np_; ++t) {
          w += displacement_cubic_kernel(px[size_t(i) * K + t], c);
          g += displacement_cubic_kernel_grad(px[size_t(i) * K + t], c);
        }
        out[4 * i] = c.vol * g.x;
        out[4 * i + 1] = c.vol * g.y;
        out[4 * i + 2] = c.vol * g.z;
        out[4 * i + 3] = c.vol * w;
      } else if (what == 1) {
        T dens = c.w0;
        for (unsigned t = 0; t < nb; ++t) dens += displacement_cubic_kernel(bx[size_t(i) * K + t], c);
        for (unsigned t = 0; t < np_; ++t) dens += displacement_cubic_kernel(px[size_t(i) * K + t], c);
        out[i] = dens * c.mass;
      } else {
        T sgk = 0;
        V3<T> gpi{0, 0, 0};
        for (unsigned t = 0; t < nb; ++t) {
          V3<T> gp = (-c.vol) * displacement_cubic_kernel_grad(bx[size_t(i) * K + t], c);
          sgk += length_sqr(gp);
          gpi -= gp;
        }
        for (unsigned t = 0; t < np_; ++t) {
          V3<T> gp = T(-1) * px[size_t(i) * K + t];
          sgk += length_sqr(gp);
          gpi -= gp;
        }
        sgk += length_sqr(gpi);
        out[i] = sgk > c.eps ? -1 / sgk : T(0);
      }
    }
  }
  // SolverI::step with pellets (solver_i.hpp:79-330, VolumeMethod::pellets)
  void step_i_pellets() {
    clear_force();
    for (unsigned i = 0; i < n; ++i) a[i] = c.g;
    snapshot_kinematics();
    transform_all_pellets();
    update_particle_neighbors();
    compute_viscosity_with_pellets();
    update_dt();
    for (unsigned i = 0; i < n; ++i) {
      v[i] += dt * a[i];
      last_pressure[i] = pressure[i] * static_cast<T>(0.5);
    }
    const T eps = static_cast<T>(0.01) * c.h2;
    // calculate_isph_diagonal_adv_density_with_pellets (runner.hpp:2951-3001)
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(n); ++ii) {
      unsigned i = unsigned(ii);
      V3<T> vi = v[i];
      T dens = density[i];
      T dadv = dens / c.rho0;
      T diag = 0;
      for (unsigned s = 0; s < num_neighbors[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = nb_id[o];
        V3<T> xixj = nb_x[o];
        V3<T> grad_w = displacement_cubic_kernel_grad(xixj, c);
        T dj = density[j];
        dadv += dt * c.vol * dot(vi - v[j], grad_w);
        diag += c.vol * (dens + dj) / (dens * dj) / (length_sqr(xixj) + eps) * dot(xixj, grad_w);
      }
      for (unsigned s = 0; s < num_pnb[i]; ++s) {
        size_t o = size_t(i) * c.kmax + s;
        unsigned j = pnb_id[o];
        V3<T> xixj = pnb_x[o];
        V3<T> g = c.vol * displacement_cubic_kernel_grad(xixj, c);
        dadv += dt * dot(vi - v[j], g);
        diag += (dens + c.rho0) / (dens * c.rho0) / (length_sqr(xixj) + eps) * dot(xixj, g);
      }
      aii[i] = diag * c.rho0;
      adv_density[i] = dadv;
    }
    T mean = std::numeric_limits<T>::max();
    unsigned it = 0;
    while ((mean > static_cast<T>(set.density_error_tolerance) || it < set.min_density_solve) &&
           it < set.max_density_solve) {
      extrapolate_pressure(last_pressure);
      // isph_solve_iteration_with_pellets (runner.hpp:3055-3118)
#pragma omp parallel for schedule(static)
      for (long ii = 0; ii < long(n); ++ii) {
        unsigned i = unsigned(ii);
        T dens = density[i];
        T diag = aii[i];
        T bb = 1 - adv_density[i];
        T dt2 = dt * dt;
        T denom = diag * dt2;
        const T w = static_cast<T>(2) / 3;
        T off = 0;
        for (unsigned s = 0; s < num_neighbors[i]; ++s) {
          size_t o = size_t(i) * c.kmax + s;
          unsigned j = nb_id[o];
          V3<T> xixj = nb_x[o];
          T dj = density[j];
          off -= c.vol * (dens + dj) / (dens * dj) / (length_sqr(xixj) + eps) *
                 dot(xixj, displacement_cubic_kernel_grad(xixj, c)) * last_pressure[j];
        }
        for (unsigned s = 0; s < num_pnb[i]; ++s) {
          size_t o = size_t(i) * c.kmax + s;
          unsigned j = pnb_id[o];
          V3<T> xixj = pnb_x[o];
          T dj = c.rho0;
          off -= c.vol * (dens + dj) / (dens * dj) / (length_sqr(xixj) + eps) *
                 dot(xixj, displacement_cubic_kernel_grad(xixj, c)) * last_pressure[j];
        }
        off *= c.rho0;
        T p = 0;
        if (std::fabs(denom) > static_cast<T>(1.0e-9))
          p = std::max((1 - w) * last_pressure[i] + w * (bb / denom - off / diag), static_cast<T>(0));
        density_err[i] = p == 0 ? 0 : (denom * p + off * dt2 - bb);
        pressure[i] = p;
      }
      for (unsigned i = 0; i < n; ++i) last_pressure[i] = pressure[i];
      mean = static_cast<T>(sum_of(density_err) / double(n));
      ++it;
    }
    rep.num_density_solve = it;
    rep.mean_density_error = mean;
    extrapolate_pressure(pressure);
    pressure_accels_and_integrate();
  }

  // ---- volume map construction (runner.hpp:1660-1694) -------------------------
  void build_volume_map(unsigned b, unsigned begin, unsigned end) {
    Boundary<T>& B = bnd[b];
    end = std::min(end, B.geom.num_nodes);
    if (B.is_map()) {
      // update_volume_field from the distance map (runner.hpp:1616-1658)
      const double dmin[3] = {double(B.geom.dmin.x), double(B.geom.dmin.y), double(B.geom.dmin.z)};
      const double dmax[3] = {double(B.geom.dmax.x), double(B.geom.dmax.y), double(B.geom.dmax.z)};
      const double cs[3] = {double(B.geom.cs.x), double(B.geom.cs.y), double(B.geom.cs.z)};
      std::vector<double> dist(B.distance.begin(), B.distance.end()), out(B.geom.num_nodes);
      volume_from_distance_impl<T>(double(c.h), dmin, dmax, B.geom.res, cs, double(B.sign),
                                   dist.data(), begin, end, out.data());
      for (unsigned l = begin; l < end; ++l) B.volume[l] = T(out[l]);
      return;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (long ll = long(begin); ll < long(end); ++ll) {
      unsigned l = unsigned(ll);
      V3<T> xn = node_position(B.geom, l);
      T sum = 0;
      for (int i = 0; i < kNumTheta; ++i) {
        T theta = 2 * c.pi * (i + 1) / kNumTheta;
        for (int j = 0; j < kNumPhi; ++j) {
          for (int k = 0; k < kNumR; ++k) {
            T wijk = 2 * c.pi * c.B[j] * c.C[k] / kNumTheta;
            V3<T> u{c.sinphi[j] * std::cos(theta), c.sinphi[j] * std::sin(theta),
                    c.cosphi[j]};
            V3<T> p = (c.h * c.R[k]) * u;
            T dist = B.sdf.signed_distance(xn + p);
            T dm = dist * B.sign;
            sum += wijk * (dm <= 0 ? T(1) : dist_cubic_kernel(dm, c) / c.w0);
          }
        }
      }
      B.volume[l] = sum * c.h2 * c.h;
    }
  }
};

}  // namespace orc

using namespace orc;

namespace {
struct Handle {
  int precision;
  std::unique_ptr<Solver<float>> f;
  std::unique_ptr<Solver<double>> d;
};
template <class F>
int dispatch(void* h, F fn) {
  if (!h) return -1;
  Handle* hh = static_cast<Handle*>(h);
  if (hh->precision == 4) return fn(*hh->f);
  return fn(*hh->d);
}

template <class T>
void add_boundary_impl(Solver<T>& s, const orc_boundary_desc& d) {
  Boundary<T> B{};
  B.sdf.set(d);
  B.sign = static_cast<T>(d.sign);
  // calculate_domain_and_resolution (pile.hpp:40-65)
  T margin = d.margin < 0 ? (B.sign < 0 ? s.c.h : s.c.h * 2)
                          : static_cast<T>(d.margin);
  B.geom.dmin = {B.sdf.aabb_min.x - margin, B.sdf.aabb_min.y - margin,
                 B.sdf.aabb_min.z - margin};
  B.geom.dmax = {B.sdf.aabb_max.x + margin, B.sdf.aabb_max.y + margin,
                 B.sdf.aabb_max.z + margin};
  unsigned res[3] = {d.resolution[0], d.resolution[1], d.resolution[2]};
  if (res[0] == 0 || res[1] == 0 || res[2] == 0) {
    V3<T> ext = B.geom.dmax - B.geom.dmin;
    T cw = s.c.radius;
    if (res[0] == 0) res[0] = static_cast<unsigned>(ext.x / cw);
    if (res[1] == 0) res[1] = static_cast<unsigned>(ext.y / cw);
    if (res[2] == 0) res[2] = static_cast<unsigned>(ext.z / cw);
  }
  // calculate_grid_attributes (pile.hpp:66-78)
  for (int a = 0; a < 3; ++a) B.geom.res[a] = res[a];
  V3<T> ext = B.geom.dmax - B.geom.dmin;
  B.geom.cs = {ext.x / T(res[0]), ext.y / T(res[1]), ext.z / T(res[2])};
  unsigned nv = (res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  unsigned ne = res[0] * (res[1] + 1) * (res[2] + 1) +
                (res[0] + 1) * res[1] * (res[2] + 1) +
                (res[0] + 1) * (res[1] + 1) * res[2];
  B.geom.num_nodes = nv + 2 * ne;
  B.volume.assign(B.geom.num_nodes, 0);
  B.mass = static_cast<T>(d.mass);
  B.inertia = {T(d.inertia[0]), T(d.inertia[1]), T(d.inertia[2])};
  B.x = {T(d.x[0]), T(d.x[1]), T(d.x[2])};
  B.q = {T(d.q[0]), T(d.q[1]), T(d.q[2]), T(d.q[3])};
  B.v = {0, 0, 0};
  B.omega = {0, 0, 0};
  B.force = {0, 0, 0};
  B.torque = {0, 0, 0};
  B.a = {0, 0, 0};
  s.bnd.push_back(B);
  s.realloc_boundary();
}

template <class T, class E>
int copy_vec(std::vector<E>& v, size_t count, const void* src, bool in,
             void* dst) {
  if (in) {
    std::memcpy(v.data(), src, count * sizeof(E));
  } else {
    std::memcpy(dst, v.data(), count * sizeof(E));
  }
  return 0;
}

template <class T>
int array_io(Solver<T>& s, const char* name, const void* src, void* dst) {
  bool in = src != nullptr;
  std::string nm(name);
  size_t N = s.max_n;
  size_t NT = s.nt();
  size_t BN = s.bnd.size() * N;
#define ALV_ORC_ARR(str, vec, cnt)                     \
  if (nm == str) return copy_vec<T>(vec, cnt, src, in, dst);
  ALV_ORC_ARR("x", s.x, NT)
  ALV_ORC_ARR("v", s.v, NT)
  ALV_ORC_ARR("a", s.a, N)
  ALV_ORC_ARR("density", s.density, N)
  ALV_ORC_ARR("dfsph_factor", s.dfsph_factor, N)
  ALV_ORC_ARR("kappa", s.kappa, N)
  ALV_ORC_ARR("kappa_v", s.kappa_v, N)
  ALV_ORC_ARR("density_adv", s.density_adv, N)
  ALV_ORC_ARR("cfl_v2", s.cfl_v2, N)
  ALV_ORC_ARR("pressure", s.pressure, NT)
  ALV_ORC_ARR("last_pressure", s.last_pressure, NT)
  ALV_ORC_ARR("aii", s.aii, N)
  ALV_ORC_ARR("dii", s.dii, N)
  ALV_ORC_ARR("dij_pj", s.dij_pj, N)
  ALV_ORC_ARR("adv_density", s.adv_density, N)
  ALV_ORC_ARR("pressure_accel", s.pressure_accel, N)
  ALV_ORC_ARR("density_err", s.density_err, N)
  ALV_ORC_ARR("num_neighbors", s.num_neighbors, NT)
  ALV_ORC_ARR("num_pellet_neighbors", s.num_pnb, NT)
  ALV_ORC_ARR("boundary", s.boundary, BN)
  ALV_ORC_ARR("boundary_kernel", s.boundary_kernel, BN)
  ALV_ORC_ARR("force", s.force, BN)
  ALV_ORC_ARR("torque", s.torque, BN)
#undef ALV_ORC_ARR
  return -2;
}
}  // namespace

extern "C" {

void* orc_create(int precision, int kind, unsigned max_num_particles,
                 const orc_params* params) {
  if (!params || (precision != 4 && precision != 8)) return nullptr;
  Handle* h = new Handle();
  h->precision = precision;
  auto init = [&](auto& s) {
    s.kind = kind;
    s.max_n = max_num_particles;
    s.n = max_num_particles;
    s.c.set(*params);
    s.set = orc_settings{2e-4, 2e-4, 2e-4, 0.4, 1e-3, 1e-3, 2, 100, 1, 100, 1, 1};
    std::memset(&s.rep, 0, sizeof(s.rep));
    s.alloc();
  };
  if (precision == 4) {
    h->f.reset(new Solver<float>());
    init(*h->f);
  } else {
    h->d.reset(new Solver<double>());
    init(*h->d);
  }
  return h;
}

void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

int orc_set_settings(void* h, const orc_settings* st) {
  return dispatch(h, [&](auto& s) {
    s.set = *st;
    s.dt = static_cast<decltype(s.dt)>(st->dt);
    return 0;
  });
}

int orc_set_num_particles(void* h, unsigned n) {
  return dispatch(h, [&](auto& s) {
    if (n > s.max_n) return -3;
    s.n = n;
    return 0;
  });
}

int orc_add_boundary(void* h, const orc_boundary_desc* d) {
  return dispatch(h, [&](auto& s) {
    add_boundary_impl(s, *d);
    return 0;
  });
}

unsigned orc_num_nodes(void* h, unsigned b) {
  unsigned r = 0;
  dispatch(h, [&](auto& s) {
    if (b < s.bnd.size()) r = s.bnd[b].geom.num_nodes;
    return 0;
  });
  return r;
}

int orc_get_grid_info(void* h, unsigned b, double* dmin, double* dmax,
                      unsigned* res, double* cs) {
  return dispatch(h, [&](auto& s) {
    if (b >= s.bnd.size()) return -3;
    auto& g = s.bnd[b].geom;
    dmin[0] = g.dmin.x;
    dmin[1] = g.dmin.y;
    dmin[2] = g.dmin.z;
    dmax[0] = g.dmax.x;
    dmax[1] = g.dmax.y;
    dmax[2] = g.dmax.z;
    cs[0] = g.cs.x;
    cs[1] = g.cs.y;
    cs[2] = g.cs.z;
    for (int a = 0; a < 3; ++a) res[a] = g.res[a];
    return 0;
  });
}

int orc_build_volume_map(void* h, unsigned b, unsigned begin, unsigned end) {
  return dispatch(h, [&](auto& s) {
    if (b >= s.bnd.size()) return -3;
    s.build_volume_map(b, begin, end);
    return 0;
  });
}

int orc_set_volume_map(void* h, unsigned b, const void* nodes) {
  return dispatch(h, [&](auto& s) {
    if (b >= s.bnd.size()) return -3;
    auto& vol = s.bnd[b].volume;
    std::memcpy(vol.data(), nodes, vol.size() * sizeof(vol[0]));
    return 0;
  });
}

int orc_set_distance_map(void* h, unsigned b, const void* nodes) {
  return dispatch(h, [&](auto& s) {
    if (b >= s.bnd.size() || !s.bnd[b].is_map()) return -3;
    auto& d = s.bnd[b].distance;
    d.resize(s.bnd[b].geom.num_nodes);
    std::memcpy(d.data(), nodes, d.size() * sizeof(d[0]));
    return 0;
  });
}

int orc_get_volume_map(void* h, unsigned b, void* out) {
  return dispatch(h, [&](auto& s) {
    if (b >= s.bnd.size()) return -3;
    auto& vol = s.bnd[b].volume;
    std::memcpy(out, vol.data(), vol.size() * sizeof(vol[0]));
    return 0;
  });
}

int orc_set_rigid(void* h, unsigned b, const double* x, const double* v,
                  const double* q, const double* w) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    if (b >= s.bnd.size()) return -3;
    auto& B = s.bnd[b];
    if (x) B.x = {T(x[0]), T(x[1]), T(x[2])};
    if (v) B.v = {T(v[0]), T(v[1]), T(v[2])};
    if (q) B.q = {T(q[0]), T(q[1]), T(q[2]), T(q[3])};
    if (w) B.omega = {T(w[0]), T(w[1]), T(w[2])};
    return 0;
  });
}

}  // extern "C"

namespace {
template <class T>
std::vector<V3<T>> v3s(const double* a, size_t n) {
  std::vector<V3<T>> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = {T(a[3 * i]), T(a[3 * i + 1]), T(a[3 * i + 2])};
  return out;
}
template <class T>
std::vector<V4<T>> v4s(const double* a, size_t n) {
  std::vector<V4<T>> out(n);
  for (size_t i = 0; i < n; ++i)
    out[i] = {T(a[4 * i]), T(a[4 * i + 1]), T(a[4 * i + 2]), T(a[4 * i + 3])};
  return out;
}
}  // namespace

extern "C" {

int orc_sample_neighbors(void* h, const double* sample_x, unsigned ns, unsigned* ids,
                         double* xixj, unsigned* counts) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    if (s.grid_x.size() != (s.pellet_mode ? s.nt() : size_t(s.n))) return -4;  // no grid yet
    auto sx = v3s<T>(sample_x, ns);
    std::vector<V3<T>> d(size_t(ns) * s.c.kmax, V3<T>{0, 0, 0});
    s.sample_neighbors(sx.data(), ns, ids, d.data(), counts);
    for (size_t k = 0; k < d.size(); ++k) {
      xixj[3 * k] = d[k].x;
      xixj[3 * k + 1] = d[k].y;
      xixj[3 * k + 2] = d[k].z;
    }
    return 0;
  });
}

int orc_sample_boundary(void* h, const double* sample_x, unsigned ns, double* rec, double* ker) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    auto sx = v3s<T>(sample_x, ns);
    for (unsigned b = 0; b < s.bnd.size(); ++b)
      for (unsigned i = 0; i < ns; ++i) {
        V4<T> r, k;
        s.point_boundary(b, sx[i], r, k);
        double* ro = rec + 4 * (size_t(b) * ns + i);
        double* ko = ker + 4 * (size_t(b) * ns + i);
        ro[0] = r.x; ro[1] = r.y; ro[2] = r.z; ro[3] = r.w;
        ko[0] = k.x; ko[1] = k.y; ko[2] = k.z; ko[3] = k.w;
      }
    return 0;
  });
}

int orc_sample_fluid(void* h, const unsigned* ids, const double* xixj, const unsigned* counts,
                     unsigned ns, const double* quantity, unsigned dim, double* out) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    auto d = v3s<T>(xixj, size_t(ns) * s.c.kmax);
    std::vector<T> q(size_t(s.n) * dim), o(size_t(ns) * dim);
    for (size_t k = 0; k < q.size(); ++k) q[k] = T(quantity[k]);
    s.sample_fluid(ids, d.data(), counts, ns, q.data(), dim, o.data());
    for (size_t k = 0; k < o.size(); ++k) out[k] = o[k];
    return 0;
  });
}

int orc_sample_density(void* h, const double* xixj, const unsigned* counts, unsigned ns,
                       const double* ker, unsigned nb, double* out) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    auto d = v3s<T>(xixj, size_t(ns) * s.c.kmax);
    auto k = nb ? v4s<T>(ker, size_t(nb) * ns) : std::vector<V4<T>>();
    std::vector<T> o(ns);
    s.sample_density(d.data(), counts, ns, k.data(), nb, o.data());
    for (unsigned i = 0; i < ns; ++i) out[i] = o[i];
    return 0;
  });
}

int orc_sample_velocity(void* h, const unsigned* ids, const double* xixj, const unsigned* counts,
                        unsigned ns, const double* sample_v, const double* rec, const double* ker,
                        unsigned nb, const double* rigid_x, const double* rigid_v,
                        const double* rigid_omega, double* out) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    auto d = v3s<T>(xixj, size_t(ns) * s.c.kmax);
    std::vector<V3<T>> sv = sample_v ? v3s<T>(sample_v, ns) : std::vector<V3<T>>();
    auto r = nb ? v4s<T>(rec, size_t(nb) * ns) : std::vector<V4<T>>();
    auto k = nb ? v4s<T>(ker, size_t(nb) * ns) : std::vector<V4<T>>();
    auto rx = nb ? v3s<T>(rigid_x, nb) : std::vector<V3<T>>();
    auto rv = nb ? v3s<T>(rigid_v, nb) : std::vector<V3<T>>();
    auto rw = nb ? v3s<T>(rigid_omega, nb) : std::vector<V3<T>>();
    std::vector<V3<T>> o(ns);
    s.sample_velocity(ids, d.data(), counts, ns, sample_v ? sv.data() : nullptr, r.data(),
                      k.data(), nb, rx.data(), rv.data(), rw.data(), o.data());
    for (unsigned i = 0; i < ns; ++i) {
      out[3 * i] = o[i].x;
      out[3 * i + 1] = o[i].y;
      out[3 * i + 2] = o[i].z;
    }
    return 0;
  });
}

int orc_set_collision(void* h, unsigned b, double restitution, double friction,
                      const double* xyz, unsigned n) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    if (b >= s.bnd.size()) return -3;
    auto& B = s.bnd[b];
    B.restitution = T(restitution);
    B.friction = T(friction);
    B.coll.resize(n);
    for (unsigned k = 0; k < n; ++k)
      B.coll[k] = {T(xyz[3 * k]), T(xyz[3 * k + 1]), T(xyz[3 * k + 2])};
    return 0;
  });
}

int orc_set_contact_params(void* h, double contact_tolerance, unsigned max_num_contacts) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    s.contact_tolerance = T(contact_tolerance);
    s.max_num_contacts = max_num_contacts;
    return 0;
  });
}

int orc_find_contacts(void* h, int i, int j, unsigned* num_found) {
  return dispatch(h, [&](auto& s) {
    if (i >= int(s.bnd.size()) || j >= int(s.bnd.size()) || (i < 0 && j >= 0)) return -3;
    s.find_contacts(i, j);
    if (num_found) *num_found = s.num_found;
    return 0;
  });
}

int orc_solve_contacts(void* h) {
  return dispatch(h, [&](auto& s) {
    s.solve_contacts();
    return 0;
  });
}

int orc_get_contacts(void* h, double* out, unsigned capacity, unsigned* n) {
  return dispatch(h, [&](auto& s) {
    unsigned k = std::min(capacity, unsigned(s.contacts.size()));
    for (unsigned c = 0; c < k; ++c) {
      const auto& C = s.contacts[c];
      double* o = out + size_t(c) * ORC_CONTACT_FIELDS;
      o[0] = C.i;
      o[1] = C.j;
      const decltype(C.n)* vs[8] = {&C.cp_i, &C.cp_j, &C.n, &C.t, &C.iiwi_diag,
                                    &C.iiwi_off_diag, &C.iiwj_diag, &C.iiwj_off_diag};
      for (int f = 0; f < 8; ++f) {
        o[2 + 3 * f] = vs[f]->x;
        o[3 + 3 * f] = vs[f]->y;
        o[4 + 3 * f] = vs[f]->z;
      }
      o[26] = C.friction;
      o[27] = C.nkninv;
      o[28] = C.pmax;
      o[29] = C.goalu;
      o[30] = C.impulse_sum;
    }
    if (n) *n = k;
    return 0;
  });
}

int orc_get_rigid(void* h, unsigned b, double* x, double* v, double* q,
                  double* w, double* f, double* tq) {
  return dispatch(h, [&](auto& s) {
    if (b >= s.bnd.size()) return -3;
    auto& B = s.bnd[b];
    if (x) {
      x[0] = B.x.x;
      x[1] = B.x.y;
      x[2] = B.x.z;
    }
    if (v) {
      v[0] = B.v.x;
      v[1] = B.v.y;
      v[2] = B.v.z;
    }
    if (q) {
      q[0] = B.q.x;
      q[1] = B.q.y;
      q[2] = B.q.z;
      q[3] = B.q.w;
    }
    if (w) {
      w[0] = B.omega.x;
      w[1] = B.omega.y;
      w[2] = B.omega.z;
    }
    if (f) {
      f[0] = B.force.x;
      f[1] = B.force.y;
      f[2] = B.force.z;
    }
    if (tq) {
      tq[0] = B.torque.x;
      tq[1] = B.torque.y;
      tq[2] = B.torque.z;
    }
    return 0;
  });
}

int orc_set_gravity_pile(void* h, const double* g) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    s.pile_gravity = {T(g[0]), T(g[1]), T(g[2])};
    return 0;
  });
}

int orc_set_array(void* h, const char* name, const void* data) {
  return dispatch(h, [&](auto& s) { return array_io(s, name, data, nullptr); });
}

int orc_get_array(void* h, const char* name, void* out) {
  return dispatch(h, [&](auto& s) { return array_io(s, name, nullptr, out); });
}

int orc_compute_all_boundaries(void* h) {
  return dispatch(h, [&](auto& s) {
    s.compute_all_boundaries();
    return 0;
  });
}

int orc_update_particle_neighbors(void* h) {
  return dispatch(h, [&](auto& s) {
    s.update_particle_neighbors();
    return 0;
  });
}

int orc_compute_density_change(void* h) {
  return dispatch(h, [&](auto& s) {
    // compute_velocity_of_density_change (runner.hpp:3297-3314), sum only
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(s.n); ++i) s.density_adv[i] = s.density_change(unsigned(i));
    return 0;
  });
}

int orc_integrate_velocity(void* h) {
  return dispatch(h, [&](auto& s) {
    s.integrate_velocity();
    return 0;
  });
}

int orc_plant(int code) {
  orc::g_plant = code;
  return 0;
}

int orc_compute_dfsph_factor(void* h) {
  return dispatch(h, [&](auto& s) {
    s.compute_dfsph_factor();
    return 0;
  });
}

int orc_apply_stiffness(void* h, int divergence_form, const double* k) {
  return dispatch(h, [&](auto& s) {
    using T = typename std::remove_reference<decltype(s.dt)>::type;
    std::vector<T> kk(s.n);
    for (unsigned i = 0; i < s.n; ++i) kk[i] = static_cast<T>(k[i]);
    std::vector<orc::V3<T>> dvs(s.n);
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < long(s.n); ++ii) {
      unsigned i = unsigned(ii);
      dvs[i] = s.kappa_dv(i, kk[i], s.dt, divergence_form != 0, [&](unsigned j) { return kk[j]; });
    }
    for (unsigned i = 0; i < s.n; ++i) s.v[i] += dvs[i];
    return 0;
  });
}

int orc_compute_pressure_accels(void* h) {
  return dispatch(h, [&](auto& s) {
    s.compute_pressure_accels();
    return 0;
  });
}

int orc_get_neighbors(void* h, unsigned* ids, void* xij) {
  return dispatch(h, [&](auto& s) {
    size_t cnt = size_t(s.max_n) * s.c.kmax;
    if (ids) std::memcpy(ids, s.nb_id.data(), cnt * sizeof(unsigned));
    if (xij) std::memcpy(xij, s.nb_x.data(), cnt * sizeof(s.nb_x[0]));
    return 0;
  });
}

int orc_enable_pellets(void* h, unsigned max_num_pellets) {
  return dispatch(h, [&](auto& s) {
    s.pellet_mode = true;
    s.max_p = max_num_pellets;
    s.alloc();
    return 0;
  });
}

int orc_refresh_pellets(void* h, unsigned* num_pellets) {
  return dispatch(h, [&](auto& s) {
    int rc = s.refresh_pellets();
    if (num_pellets) *num_pellets = s.np;
    return rc;
  });
}

int orc_hint_identical_sequence(void* h, unsigned begin, unsigned end) {
  return dispatch(h, [&](auto& s) {
    s.ident.push_back({begin, end});
    return 0;
  });
}

int orc_transform_pellets(void* h) {
  return dispatch(h, [&](auto& s) {
    if (!s.pellet_mode || s.refresh_pellets() != 0) return -4;
    s.snapshot_kinematics();
    s.transform_all_pellets();
    return 0;
  });
}

int orc_find_contacts_pellets(void* h, unsigned* num_found) {
  return dispatch(h, [&](auto& s) {
    if (!s.pellet_mode || s.refresh_pellets() != 0) return -4;
    s.snapshot_kinematics();
    s.transform_all_pellets();
    s.find_contacts_pellets();
    if (num_found) *num_found = s.num_found;
    return 0;
  });
}

int orc_extrapolate_pressure(void* h, int last) {
  return dispatch(h, [&](auto& s) {
    if (!s.pellet_mode) return -4;
    s.extrapolate_pressure(last ? s.last_pressure : s.pressure);
    return 0;
  });
}

int orc_sample_bead_pellet(void* h, const double* sample_x, unsigned ns, unsigned offset,
                           unsigned max_beads, unsigned* bead_ids, double* bead_xixj,
                           unsigned* bead_counts, unsigned* pel_ids, double* pel_xixj,
                           unsigned* pel_counts) {
  return dispatch(h, [&](auto& s) {
    using T = std::remove_cv_t<std::remove_reference_t<decltype(s.c.h)>>;
    const size_t rows = size_t(offset) + ns, K = s.c.kmax;
    std::vector<V3<T>> sx(rows), bx(rows * K, V3<T>{0, 0, 0}), px(rows * K, V3<T>{0, 0, 0});
    for (size_t i = 0; i < rows; ++i)
      sx[i] = {T(sample_x[3 * i]), T(sample_x[3 * i + 1]), T(sample_x[3 * i + 2])};
    s.sample_bead_pellet(sx.data(), ns, offset, max_beads, bead_ids, bx.data(), bead_counts,
                         pel_ids, px.data(), pel_counts);
    for (size_t o = 0; o < rows * K; ++o) {
      bead_xixj[3 * o] = bx[o].x;
      bead_xixj[3 * o + 1] = bx[o].y;
      bead_xixj[3 * o + 2] = bx[o].z;
      pel_xixj[3 * o] = px[o].x;
      pel_xixj[3 * o + 1] = px[o].y;
      pel_xixj[3 * o + 2] = px[o].z;
    }
    return 0;
  });
}

int orc_sample_with_pellets(void* h, int what, const unsigned* ids, const double* xixj,
                            const unsigned* counts, const unsigned* pids, const double* pxixj,
                            const unsigned* pcounts, unsigned ns, const double* density,
                            const double* v, const double* sample_v, double* out) {
  return dispatch(h, [&](auto& s) {
    using T = std::remove_cv_t<std::remove_reference_t<decltype(s.c.h)>>;
    const unsigned K = s.c.kmax;
    const auto& c = s.c;
    for (unsigned i = 0; i < ns; ++i) {
      const unsigned nb = std::min(counts[i], K), np_ = std::min(pcounts[i], K);
      auto d3 = [&](const double* a, size_t o) { return V3<T>{T(a[3 * o]), T(a[3 * o + 1]), T(a[3 * o + 2])}; };
      if (what == 0) {
        // sample_position_density_with_pellets (runner.hpp:2122-2147)
        T dens = 0;
        for (unsigned t = 0; t < nb; ++t) dens += displacement_cubic_kernel(d3(xixj, size_t(i) * K + t), c);
        for (unsigned t = 0; t < np_; ++t) dens += displacement_cubic_kernel(d3(pxixj, size_t(i) * K + t), c);
        out[i] = double(dens * c.mass);
        continue;
      }
      // compute_sample_velocity / vorticity_with_pellets (runner.hpp:4242-4359)
      const V3<T> vi = what == 2 ? d3(sample_v, i) : V3<T>{0, 0, 0};
      V3<T> r{0, 0, 0};
      for (unsigned t = 0; t < nb; ++t) {
        const size_t o = size_t(i) * K + t;
        const unsigned j = ids[o];
        const V3<T> d = d3(xixj, o);
        if (what == 1)
          r += (c.mass / T(density[j]) * displacement_cubic_kernel(d, c)) * d3(v, j);
        else
          r += (c.mass / T(density[j])) * cross(d3(v, j) - vi, displacement_cubic_kernel_grad(d, c));
      }
      for (unsigned t = 0; t < np_; ++t) {
        const size_t o = size_t(i) * K + t;
        const unsigned j = pids[o];
        const V3<T> d = d3(pxixj, o);
        if (what == 1)
          r += (c.vol * displacement_cubic_kernel(d, c)) * d3(v, j);
        else
          r += c.vol * cross(d3(v, j) - vi, displacement_cubic_kernel_grad(d, c));
      }
      out[3 * i] = r.x;
      out[3 * i + 1] = r.y;
      out[3 * i + 2] = r.z;
    }
    return 0;
  });
}

int orc_pellet_sums(void* h, int what, unsigned K, unsigned n, const double* bead_xixj,
                    const unsigned* bead_counts, const double* pel_xixj,
                    const unsigned* pel_counts, double* out) {
  return dispatch(h, [&](auto& s) {
    using T = std::remove_cv_t<std::remove_reference_t<decltype(s.c.h)>>;
    std::vector<V3<T>> bx(size_t(n) * K), px(size_t(n) * K);
    for (size_t o = 0; o < size_t(n) * K; ++o) {
      if (bead_xixj) bx[o] = {T(bead_xixj[3 * o]), T(bead_xixj[3 * o + 1]), T(bead_xixj[3 * o + 2])};
      px[o] = {T(pel_xixj[3 * o]), T(pel_xixj[3 * o + 1]), T(pel_xixj[3 * o + 2])};
    }
    std::vector<T> res(size_t(n) * (what == 0 ? 4 : 1));
    s.pellet_sums(what, K, n, bx.data(), bead_xixj ? bead_counts : nullptr, px.data(), pel_counts,
                  res.data());
    for (size_t k = 0; k < res.size(); ++k) out[k] = double(res[k]);
    return 0;
  });
}

int orc_get_lists(void* h, int pellet_list, unsigned* ids, void* xij) {
  return dispatch(h, [&](auto& s) {
    size_t cnt = s.nt() * s.c.kmax;
    if (pellet_list && !s.pellet_mode) return -4;
    const auto& I = pellet_list ? s.pnb_id : s.nb_id;
    const auto& X = pellet_list ? s.pnb_x : s.nb_x;
    if (ids) std::memcpy(ids, I.data(), cnt * sizeof(unsigned));
    if (xij) std::memcpy(xij, X.data(), cnt * sizeof(X[0]));
    return 0;
  });
}

int orc_get_grid(void* h, unsigned* cell_count, unsigned* cell_start,
                 unsigned* anomaly) {
  return dispatch(h, [&](auto& s) {
    if (cell_count)
      std::memcpy(cell_count, s.cell_count.data(),
                  s.cell_count.size() * sizeof(unsigned));
    if (cell_start)
      std::memcpy(cell_start, s.cell_start.data(),
                  s.cell_start.size() * sizeof(unsigned));
    if (anomaly) std::memcpy(anomaly, s.anomaly, 3 * sizeof(unsigned));
    return 0;
  });
}

int orc_step(void* h) {
  return dispatch(h, [&](auto& s) {
    if (s.pellet_mode && (s.kind != 2 || s.refresh_pellets() != 0)) return -4;
    if (s.kind == 0)
      s.step_df();
    else if (s.kind == 2)
      s.step_i();
    else
      s.step_ii();
    return 0;
  });
}

int orc_get_report(void* h, orc_report* r) {
  return dispatch(h, [&](auto& s) {
    *r = s.rep;
    r->t = s.t;
    r->dt = s.dt;
    return 0;
  });
}

int orc_peirce_constants(double* cosphi, double* sinphi, double* b, double* r,
                         double* c) {
  const PeirceTable& t = peirce();
  std::memcpy(cosphi, t.cosphi, sizeof(t.cosphi));
  std::memcpy(sinphi, t.sinphi, sizeof(t.sinphi));
  std::memcpy(b, t.b, sizeof(t.b));
  std::memcpy(r, t.r, sizeof(t.r));
  std::memcpy(c, t.c, sizeof(t.c));
  return 0;
}

}  // extern "C"
template <class T>
static void cubic_kernel_impl(double h, size_t n, const double* d,
                              double* out) {
  orc_params p{};
  p.kernel_radius = h;
  p.particle_mass = 1;
  p.density0 = 1;
  Consts<T> c;
  c.set(p);
  for (size_t i = 0; i < n; ++i) {
    V3<T> dd{T(d[3 * i]), T(d[3 * i + 1]), T(d[3 * i + 2])};
    V3<T> g = displacement_cubic_kernel_grad(dd, c);
    out[4 * i] = g.x;
    out[4 * i + 1] = g.y;
    out[4 * i + 2] = g.z;
    out[4 * i + 3] = displacement_cubic_kernel(dd, c);
  }
}

extern "C" {
int orc_cubic_kernel(int precision, double h, size_t n, const double* d,
                     double* out) {
  if (precision == 4)
    cubic_kernel_impl<float>(h, n, d, out);
  else
    cubic_kernel_impl<double>(h, n, d, out);
  return 0;
}

int orc_shape_functions(const double* xi, double* N, double* dN) {
  V3<double> x{xi[0], xi[1], xi[2]};
  double dx[32], dy[32], dz[32];
  shape_functions<double>(x, N, dx, dy, dz);
  if (dN) {
    for (int j = 0; j < 32; ++j) {
      dN[3 * j] = dx[j];
      dN[3 * j + 1] = dy[j];
      dN[3 * j + 2] = dz[j];
    }
  }
  return 0;
}

int orc_node_positions(const double* dmin, const unsigned* res, const double* cs,
                       double* out) {
  GridGeom<double> g{};
  g.dmin = {dmin[0], dmin[1], dmin[2]};
  g.cs = {cs[0], cs[1], cs[2]};
  for (int a = 0; a < 3; ++a) g.res[a] = res[a];
  const unsigned nv = (res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  const unsigned ne = res[0] * (res[1] + 1) * (res[2] + 1) + (res[0] + 1) * res[1] * (res[2] + 1) +
                      (res[0] + 1) * (res[1] + 1) * res[2];
  g.num_nodes = nv + 2 * ne;
  for (unsigned l = 0; l < g.num_nodes; ++l) {
    V3<double> x = node_position(g, l);
    out[3 * size_t(l)] = x.x;
    out[3 * size_t(l) + 1] = x.y;
    out[3 * size_t(l) + 2] = x.z;
  }
  return 0;
}

int orc_node_position(const double* dmin, const unsigned* res,
                      const double* cs, unsigned l, double* out) {
  GridGeom<double> g{};
  g.dmin = {dmin[0], dmin[1], dmin[2]};
  g.cs = {cs[0], cs[1], cs[2]};
  for (int a = 0; a < 3; ++a) g.res[a] = res[a];
  V3<double> x = node_position(g, l);
  out[0] = x.x;
  out[1] = x.y;
  out[2] = x.z;
  return 0;
}

int orc_cell_nodes(const unsigned* res, const int* ipos, unsigned* out) {
  cell_nodes(res, ipos, out);
  return 0;
}

}  // extern "C"
template <class T>
static void volume_from_distance_impl(double h, const double* dmin,
                                      const double* dmax, const unsigned* res,
                                      const double* cs, double sign,
                                      const double* dist, unsigned begin,
                                      unsigned end, double* out) {
  // update_volume_field, grid variant (runner.hpp:1616-1658)
  orc_params p{};
  p.kernel_radius = h;
  p.particle_mass = 1;
  p.density0 = 1;
  Consts<T> c;
  c.set(p);
  GridGeom<T> g{};
  g.dmin = {T(dmin[0]), T(dmin[1]), T(dmin[2])};
  g.dmax = {T(dmax[0]), T(dmax[1]), T(dmax[2])};
  g.cs = {T(cs[0]), T(cs[1]), T(cs[2])};
  for (int a = 0; a < 3; ++a) g.res[a] = res[a];
  unsigned nv = (res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  unsigned ne = res[0] * (res[1] + 1) * (res[2] + 1) +
                (res[0] + 1) * res[1] * (res[2] + 1) +
                (res[0] + 1) * (res[1] + 1) * res[2];
  g.num_nodes = nv + 2 * ne;
  std::vector<T> nodes(g.num_nodes);
  for (unsigned i = 0; i < g.num_nodes; ++i) nodes[i] = T(dist[i]);
  T sg = T(sign);
#pragma omp parallel for schedule(dynamic, 16)
  for (long ll = long(begin); ll < long(end); ++ll) {
    unsigned l = unsigned(ll);
    V3<T> xn = node_position(g, l);
    T sum = 0;
    const T wi = 2 * c.pi / kNumTheta;
    for (int i = 0; i < kNumTheta; ++i) {
      T theta = 2 * c.pi * (i + 1) / kNumTheta;
      T ct = std::cos(theta), st = std::sin(theta);
      for (int j = 0; j < kNumPhi; ++j) {
        V3<T> u{c.sinphi[j] * ct, c.sinphi[j] * st, c.cosphi[j]};
        T wij = wi * c.B[j];
        for (int k = 0; k < kNumR; ++k) {
          T wijk = wij * c.C[k];
          V3<T> pp = (c.h * c.R[k]) * u;
          int ip[3];
          V3<T> inner;
          T d = std::numeric_limits<T>::max();
          if (resolve(g, xn + pp, ip, inner)) {
            T N[32];
            unsigned cells[32];
            shape_functions<T>(inner, N, nullptr, nullptr, nullptr);
            cell_nodes(g.res, ip, cells);
            d = interpolate(nodes, cells, N);
          }
          T dm = d * sg;
          sum += wijk * (dm <= 0 ? T(1) : dist_cubic_kernel(dm, c) / c.w0);
        }
      }
    }
    out[l - begin] = double(sum * c.h2 * c.h);
  }
}

extern "C" {
int orc_volume_from_distance_map(int precision, double h, const double* dmin,
                                 const double* dmax, const unsigned* res,
                                 const double* cs, double sign,
                                 const double* dist, unsigned begin,
                                 unsigned end, double* out) {
  if (precision == 4)
    volume_from_distance_impl<float>(h, dmin, dmax, res, cs, sign, dist, begin,
                                     end, out);
  else
    volume_from_distance_impl<double>(h, dmin, dmax, res, cs, sign, dist,
                                      begin, end, out);
  return 0;
}

int orc_map_value_gradient(const double* dmin, const double* dmax, const unsigned* res,
                           const double* cs, const double* nodes, size_t n,
                           const double* points, double* value, double* gradient,
                           int* inside) {
  GridGeom<double> g{};
  g.dmin = {dmin[0], dmin[1], dmin[2]};
  g.dmax = {dmax[0], dmax[1], dmax[2]};
  g.cs = {cs[0], cs[1], cs[2]};
  for (int a = 0; a < 3; ++a) g.res[a] = res[a];
  for (size_t i = 0; i < n; ++i) {
    V3<double> x{points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    int ip[3];
    V3<double> inner, gr{0, 0, 0};
    double v = 0;
    inside[i] = resolve(g, x, ip, inner) ? 1 : 0;
    if (inside[i]) {
      double N[32], dNx[32], dNy[32], dNz[32];
      unsigned cells[32];
      shape_functions<double>(inner, N, dNx, dNy, dNz);
      cell_nodes(g.res, ip, cells);
      v = interpolate_and_derive(nodes, cells, N, dNx, dNy, dNz, g.cs, gr);
    }
    value[i] = v;
    gradient[3 * i] = gr.x;
    gradient[3 * i + 1] = gr.y;
    gradient[3 * i + 2] = gr.z;
  }
  return 0;
}

double orc_sum_f32(const float* v, size_t n) {
  double s = 0;
  for (size_t i = 0; i < n; ++i) s += v[i];
  return s;
}

}  // extern "C"
