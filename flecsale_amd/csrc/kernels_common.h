// Device arithmetic shared by the kernels: EOS, Euler fluxes, geometry, QR.
// Every function keeps the reference's floating-point operation order (the
// library is compiled with -ffp-contract=off): the limiter threshold, the
// HLLE branches and the symmetry QR flip on one-ulp differences, and the
// results are held bit-identical to the CPU restatement of the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace fsale {

// std::min / std::max semantics
__host__ __device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline double smax(double a, double b) { return (a < b) ? b : a; }

// One v_min_f64 / v_max_f64 each.  On ordered operands they return the value
// smin / smax return; only the sign of a zero result can differ (+0 vs -0),
// and that never changes a sum, a product's magnitude or a comparison.  A
// compare and two selects per smin, or fmin's operand canonicalisation,
// cost three instructions where these cost one.
__device__ __forceinline__ double hmin(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double hmax(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Element access by byte offset for entity arrays (cells, faces, vertices):
// batch capacities are held below 2^28 entities, so an offset of an 8-byte
// element fits in 32 bits.  Formed that way, the load or store takes the
// array's base from scalar registers plus one 32-bit offset per lane, instead
// of a 64-bit address computed per lane per array.  For k-major tables pass
// the row base (a scalar) and the lane's offset within the row.
template <class T>
__device__ __forceinline__ T ld(const T *a, uint32_t off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(a) + off);
}
template <class T>
__device__ __forceinline__ void st(T *a, uint32_t off, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(a) + off) = v;
}
// the same with the non-temporal hint: streamed data that no other lane or
// launch reads again soon, kept from displacing what is reused in the caches
template <class T>
__device__ __forceinline__ T ld_nt(const T *a, uint32_t off) {
  return __builtin_nontemporal_load(reinterpret_cast<const T *>(reinterpret_cast<const char *>(a) + off));
}
template <class T>
__device__ __forceinline__ void st_nt(T *a, uint32_t off, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(a) + off));
}
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_nt(const double2 *a, uint32_t off) {
  const dvec2 w = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(reinterpret_cast<const char *>(a) + off));
  return make_double2(w.x, w.y);
}
__device__ __forceinline__ void st_nt(double2 *a, uint32_t off, double2 v) {
  dvec2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<dvec2 *>(reinterpret_cast<char *>(a) + off));
}

// ----------------------------------------------------------------- EOS
// flecsale/eos/ideal_gas.h:178-226
__host__ __device__ inline double eos_e_dp(double g, double d, double p) { return p / (d * (g - 1.0)); }
__host__ __device__ inline double eos_p_de(double g, double d, double e) { return (g - 1.0) * d * e; }
__host__ __device__ inline double eos_a_de(double g, double e) { return sqrt(g * (g - 1.0) * e); }
__host__ __device__ inline double eos_T_de(double cv, double e) { return e / cv; }

// ------------------------------------------------------------ geometry
// 2D polygon: shoelace about the origin; min_length = min vertex distance.
__host__ __device__ inline void cell_geom_2d(int n, const double *px, const double *py, double &cx,
                                             double &cy, double &vol, double &minlen) {
  double a2 = 0.0, sx = 0.0, sy = 0.0;
  for (int i = 0; i < n; ++i) {
    const int j = (i + 1 == n) ? 0 : i + 1;
    const double cr = px[i] * py[j] - px[j] * py[i];
    a2 += cr;
    sx += (px[i] + px[j]) * cr;
    sy += (py[i] + py[j]) * cr;
  }
  const double area = 0.5 * a2;
  // moments times the reciprocal of 3 * (twice the area): the form whose
  // results print digit for digit as the shock_box_2d regression golden
  const double r = 1 / (3.0 * a2);
  cx = sx * r;
  cy = sy * r;
  vol = area;
  double ml = DBL_MAX;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double dx = px[j] - px[i], dy = py[j] - py[i];
      ml = smin(ml, sqrt(dx * dx + dy * dy));
    }
  minlen = ml;
}

// 2D face a->b: midpoint, unit normal (t_y, -t_x)/|t| (outward of the cell
// that traverses a->b counter-clockwise), length.
__host__ __device__ inline void face_geom_2d(double ax, double ay, double bx, double by, double &cx,
                                             double &cy, double &nx, double &ny, double &area) {
  const double tx = bx - ax, ty = by - ay;
  const double len = sqrt(tx * tx + ty * ty);
  nx = ty / len;
  ny = -tx / len;
  area = len;
  cx = 0.5 * (ax + bx);
  cy = 0.5 * (ay + by);
}

__host__ __device__ inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// 3D quad face: 4 triangles about the vertex average.
__host__ __device__ inline void quad3_geom(const double (*q)[3], double *cen, double *nrm, double &area) {
  double fm[3];
  for (int d = 0; d < 3; ++d) fm[d] = (((q[0][d] + q[1][d]) + q[2][d]) + q[3][d]) * 0.25;
  double S[3] = {0, 0, 0}, C[3] = {0, 0, 0}, A = 0.0;
  for (int i = 0; i < 4; ++i) {
    const double *p = q[i], *r = q[(i + 1) & 3];
    double u[3], v[3], w[3];
    for (int d = 0; d < 3; ++d) {
      u[d] = p[d] - fm[d];
      v[d] = r[d] - fm[d];
    }
    cross3(u, v, w);
    for (int d = 0; d < 3; ++d) w[d] = 0.5 * w[d];
    const double a = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int d = 0; d < 3; ++d) {
      S[d] += w[d];
      C[d] += a * ((fm[d] + p[d]) + r[d]);
    }
    A += a;
  }
  const double s = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  for (int d = 0; d < 3; ++d) {
    nrm[d] = S[d] / s;
    cen[d] = C[d] / (3.0 * A);
  }
  area = s;
}

// Exodus hex8 sides, outward for a positively oriented hex.
__host__ __device__ inline int hex_face_vertex(int lf, int i) {
  // {0,1,5,4} {1,2,6,5} {2,3,7,6} {0,4,7,3} {0,3,2,1} {4,5,6,7}
  constexpr int t[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                           {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};
  return t[lf][i];
}

// hex: 24 tetrahedra (cell vertex average, face vertex average, edge).
__host__ __device__ inline void hex_geom(const double (*p)[3], double *cen, double &vol, double &minlen) {
  double cm[3] = {0, 0, 0};
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) cm[d] += p[i][d];
  for (int d = 0; d < 3; ++d) cm[d] = cm[d] * 0.125;
  double V = 0.0, C[3] = {0, 0, 0};
  for (int lf = 0; lf < 6; ++lf) {
    const double *q[4];
    for (int i = 0; i < 4; ++i) q[i] = p[hex_face_vertex(lf, i)];
    double fm[3];
    for (int d = 0; d < 3; ++d) fm[d] = (((q[0][d] + q[1][d]) + q[2][d]) + q[3][d]) * 0.25;
    for (int i = 0; i < 4; ++i) {
      const double *a = q[i], *r = q[(i + 1) & 3];
      double u[3], v[3], w[3], h[3];
      for (int d = 0; d < 3; ++d) {
        u[d] = a[d] - fm[d];
        v[d] = r[d] - fm[d];
        h[d] = fm[d] - cm[d];
      }
      cross3(u, v, w);
      const double tv = (w[0] * h[0] + w[1] * h[1] + w[2] * h[2]) / 6.0;
      V += tv;
      for (int d = 0; d < 3; ++d) C[d] += tv * (((cm[d] + fm[d]) + a[d]) + r[d]);
    }
  }
  for (int d = 0; d < 3; ++d) cen[d] = C[d] / (4.0 * V);
  vol = V;
  double ml = DBL_MAX;
  for (int i = 0; i < 8; ++i)
    for (int j = i + 1; j < 8; ++j) {
      const double dx = p[j][0] - p[i][0], dy = p[j][1] - p[i][1], dz = p[j][2] - p[i][2];
      ml = smin(ml, sqrt(dx * dx + dy * dy + dz * dz));
    }
  minlen = ml;
}

// ------------------------------------------------------- Euler equations
// Primitive state with the reference tuple slot order (d, v, p, e, T, a)
// (flecsale/eqns/euler_eqns.h:84-97).
template <int N>
struct EState {
  double d, v[N], p, e, a;
};

template <int N>
__host__ __device__ inline double dotN(const double *a, const double *b) {
  double s = a[0] * b[0];
  for (int i = 1; i < N; ++i) s += a[i] * b[i];
  return s;
}

// total_energy euler_eqns.h:166-173
template <int N>
__host__ __device__ inline double e_total(const EState<N> &u) {
  return u.e + 0.5 * dotN<N>(u.v, u.v);
}

// flux euler_eqns.h:274-304
template <int N>
__host__ __device__ inline void e_flux(const EState<N> &u, const double *n, double *f) {
  const double et = e_total<N>(u);
  const double vn = dotN<N>(u.v, n);
  const double mf = u.d * vn;
  f[0] = mf;
  for (int i = 0; i < N; ++i) f[1 + i] = mf * u.v[i] + u.p * n[i];
  f[N + 1] = mf * (et + u.p / u.d);
}

// solution_delta euler_eqns.h:238-265
template <int N>
__host__ __device__ inline void e_delta(const EState<N> &l, const EState<N> &r, double *du) {
  const double el = l.d * e_total<N>(l);
  const double er = r.d * e_total<N>(r);
  du[0] = r.d - l.d;
  for (int i = 0; i < N; ++i) du[1 + i] = r.d * r.v[i] - l.d * l.v[i];
  du[N + 1] = er - el;
}

// hlle_flux flecsale/eqns/flux.h:80-106
template <int N>
__host__ __device__ inline void hlle_flux(const EState<N> &wl, const EState<N> &wr, const double *n,
                                          double *f) {
  const double vnl = dotN<N>(wl.v, n), vnr = dotN<N>(wr.v, n);
  const double ll = smin(vnl - wl.a, vnr - wr.a);
  const double lr = smax(vnl + wl.a, vnr + wr.a);
  if (ll >= 0) {
    e_flux<N>(wl, n, f);
    return;
  }
  if (lr <= 0) {
    e_flux<N>(wr, n, f);
    return;
  }
  double fl[N + 2], fr[N + 2], du[N + 2];
  e_flux<N>(wl, n, fl);
  e_flux<N>(wr, n, fr);
  const double c1 = ll * lr, c2 = lr - ll, c2inv = 1 / c2;
  e_delta<N>(wl, wr, du);
  for (int i = 0; i < N + 2; ++i) f[i] = c2inv * (lr * fl[i] - ll * fr[i] + c1 * du[i]);
}

// rusanov_flux flecsale/eqns/flux.h:52-65
template <int N>
__host__ __device__ inline void rusanov_flux(const EState<N> &wl, const EState<N> &wr, const double *n,
                                             double *f) {
  double fl[N + 2], fr[N + 2], du[N + 2];
  e_flux<N>(wl, n, fl);
  e_flux<N>(wr, n, fr);
  const double sl = wl.a + fabs(dotN<N>(wl.v, n));
  const double sr = wr.a + fabs(dotN<N>(wr.v, n));
  const double s = smax(sl, sr);
  e_delta<N>(wl, wr, du);
  const double hs = s / 2;
  for (int i = 0; i < N + 2; ++i) f[i] = (fl[i] + fr[i]) / 2 - du[i] * hs;
}

// HLLC (Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics,
// sec. 10.4) with the HLLE wave-speed bounds; not in the reference.
template <int N>
__host__ __device__ inline void hllc_flux(const EState<N> &wl, const EState<N> &wr, const double *n,
                                          double *f) {
  const double ul = dotN<N>(wl.v, n), ur = dotN<N>(wr.v, n);
  const double sL = smin(ul - wl.a, ur - wr.a);
  const double sR = smax(ul + wl.a, ur + wr.a);
  if (sL >= 0) {
    e_flux<N>(wl, n, f);
    return;
  }
  if (sR <= 0) {
    e_flux<N>(wr, n, f);
    return;
  }
  const double ml = wl.d * (sL - ul), mr = wr.d * (sR - ur);
  const double sM = (wr.p - wl.p + ml * ul - mr * ur) / (ml - mr);
  const EState<N> &w = (sM >= 0) ? wl : wr;
  const double s = (sM >= 0) ? sL : sR;
  const double un = (sM >= 0) ? ul : ur;
  double fk[N + 2];
  e_flux<N>(w, n, fk);
  const double Ek = w.d * e_total<N>(w);
  const double coef = w.d * (s - un) / (s - sM);
  double us[N + 2], uk[N + 2];
  us[0] = coef;
  for (int i = 0; i < N; ++i) us[1 + i] = coef * (w.v[i] + (sM - un) * n[i]);
  us[N + 1] = coef * (Ek / w.d + (sM - un) * (sM + w.p / (w.d * (s - un))));
  uk[0] = w.d;
  for (int i = 0; i < N; ++i) uk[1 + i] = w.d * w.v[i];
  uk[N + 1] = Ek;
  for (int i = 0; i < N + 2; ++i) f[i] = fk[i] + s * (us[i] - uk[i]);
}

// wall_flux euler_eqns.h:313-324
template <int N>
__host__ __device__ inline void wall_flux(const EState<N> &u, const double *n, double *f) {
  f[0] = 0;
  for (int i = 0; i < N; ++i) f[1 + i] = u.p * n[i];
  f[N + 1] = 0;
}

template <int N>
__host__ __device__ inline void face_flux(int kind, const EState<N> &wl, const EState<N> &wr,
                                          const double *n, double *f) {
  if (kind == 1)
    rusanov_flux<N>(wl, wr, n, f);
  else if (kind == 2)
    hllc_flux<N>(wl, wr, n, f);
  else
    hlle_flux<N>(wl, wr, n, f);
}

// ------------------------------------------------------------- QR
// Householder QR with column pivoting and the reference's back solve
// (flecsale/linalg/qr.h:44-94, detail/qr_impl.h:38-288); A row-major n x n.
template <int NMAX>
__host__ __device__ inline void qr_solve(int rows, double *A, double *B) {
  const int cols = rows;
  double v[NMAX];
  for (int i = 0; i < NMAX; ++i) v[i] = 0;
  int jpvt[NMAX];
  for (int i = 0; i < cols; ++i) jpvt[i] = i;
  for (int i = 0; i < cols; ++i) {
    double cn[NMAX];
    for (int j = 0; j < cols; ++j) {
      cn[j] = 0;
      for (int r = i; r < rows; ++r) cn[j] += A[r * cols + jpvt[j]] * A[r * cols + jpvt[j]];
    }
    int max_loc = 0;
    double mx = 0;
    for (int j = 0; j < cols; ++j)
      if (cn[j] > mx) {
        mx = cn[j];
        max_loc = j;
      }
    {
      const int t = jpvt[i];
      jpvt[i] = jpvt[max_loc];
      jpvt[max_loc] = t;
    }
    const int col = jpvt[i];
    do {
      double norm = 0;
      for (int r = i; r < rows; ++r) norm += A[r * cols + col] * A[r * cols + col];
      if (norm == 0) break;
      norm = sqrt(norm);
      v[0] = A[i * cols + col] - norm;
      for (int r = 1; r < rows - i; ++r) v[r] = A[(r + i) * cols + col];
      norm = 0;
      for (int r = 0; r < rows - i; ++r) norm += v[r] * v[r];
      if (norm == 0) break;
      norm = sqrt(norm);
      for (int r = 0; r < rows - i; ++r) v[r] *= (1.0 / norm);
    } while (0);
    const int nn = rows - i;
    double Ac[NMAX * NMAX], Bc[NMAX], hh[NMAX * NMAX];
    for (int q = 0; q < rows * cols; ++q) Ac[q] = A[q];
    for (int q = 0; q < rows; ++q) Bc[q] = B[q];
    for (int j = 0; j < nn; ++j)
      for (int r = 0; r < nn; ++r) hh[r + nn * j] = (r != j) ? -2 * v[j] * v[r] : 1 - 2 * v[j] * v[r];
    for (int k = 0; k < cols; ++k)
      for (int j = 0; j < nn; ++j) {
        double temp = 0;
        for (int r = 0; r < nn; ++r) temp += hh[r + nn * j] * Ac[(r + i) * cols + jpvt[k]];
        A[(j + i) * cols + jpvt[k]] = temp;
      }
    for (int j = 0; j < nn; ++j) {
      double sum = 0;
      for (int r = 0; r < nn; ++r) sum += hh[r + nn * j] * Bc[r + i];
      B[j + i] = sum;
    }
  }
  const double eps = DBL_EPSILON;
  double Bc[NMAX];
  for (int q = 0; q < rows; ++q) Bc[q] = B[q];
  int bottom = 0;
  for (int i = rows; i-- > 0;) {
    if (fabs(A[i * cols + jpvt[cols - 1]]) > eps) {
      bottom = i;
      break;
    }
  }
  if (bottom > cols - 1) bottom = cols - 1;
  for (int i = bottom + 1; i-- > 0;) {
    double sum = 0;
    for (int j = cols; j-- > i + 1;) sum += B[jpvt[j]] * A[i * cols + jpvt[j]];
    if (fabs(A[i * cols + jpvt[i]]) > eps) {
      const double temp = 1 / A[i * cols + jpvt[i]];
      B[jpvt[i]] = (Bc[i] - sum) * temp;
    } else {
      B[jpvt[i]] = 0;
    }
  }
}

// The same QR solve for a compile-time size, every loop unrolled so that the
// matrices stay in registers (dynamically indexed arrays would live in
// scratch or LDS).  The column pivot is applied by physically swapping
// columns instead of indirecting through jpvt: physical column j holds the
// reference's column jpvt[j], every operation reads the same values in the
// same order, and jpvt only maps the solution back (B[jpvt[i]]).
template <int NR>
__host__ __device__ inline double perm_get(const double (&x)[NR], int idx) {
  // x[idx] as the OR of each element's bits under an all-ones or all-zeros
  // mask: the select chain this replaces was turned back into an indexed
  // load from a stack copy of x (scratch memory, one round trip per read)
  unsigned long long r = 0;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    unsigned long long b;
    __builtin_memcpy(&b, &x[q], sizeof b);
    r |= b & (0ull - (unsigned long long)(idx == q));
  }
  double d;
  __builtin_memcpy(&d, &r, sizeof d);
  return d;
}

template <int NR>
__host__ __device__ inline void perm_set(double (&x)[NR], int idx, double val) {
#pragma unroll
  for (int q = 0; q < NR; ++q) x[q] = (idx == q) ? val : x[q];
}

template <int NR>
__host__ __device__ inline void qr_solve_fixed(double (&A)[NR][NR], double (&B)[NR]) {
  int jpvt[NR];
  double v[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    jpvt[q] = q;
    v[q] = 0;
  }
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    double cn[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      cn[j] = 0;
#pragma unroll
      for (int r = i; r < NR; ++r) cn[j] += A[r][j] * A[r][j];
    }
    int max_loc = 0;
    double mx = 0;
#pragma unroll
    for (int j = 0; j < NR; ++j)
      if (cn[j] > mx) {
        mx = cn[j];
        max_loc = j;
      }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      if (j != i && j == max_loc) {
        const int t = jpvt[i];
        jpvt[i] = jpvt[j];
        jpvt[j] = t;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double a = A[r][i];
          A[r][i] = A[r][j];
          A[r][j] = a;
        }
      }
    }
    {
      double norm = 0;
#pragma unroll
      for (int r = i; r < NR; ++r) norm += A[r][i] * A[r][i];
      if (norm != 0) {
        norm = sqrt(norm);
        v[0] = A[i][i] - norm;
#pragma unroll
        for (int r = 1; r < NR - i; ++r) v[r] = A[r + i][i];
        double n2 = 0;
#pragma unroll
        for (int r = 0; r < NR - i; ++r) n2 += v[r] * v[r];
        if (n2 != 0) {
          n2 = sqrt(n2);
#pragma unroll
          for (int r = 0; r < NR - i; ++r) v[r] *= (1.0 / n2);
        }
      }
    }
    double Ac[NR][NR], Bc[NR], hh[NR][NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      Bc[r] = B[r];
#pragma unroll
      for (int q = 0; q < NR; ++q) Ac[r][q] = A[r][q];
    }
#pragma unroll
    for (int j = 0; j < NR - i; ++j)
#pragma unroll
      for (int r = 0; r < NR - i; ++r) hh[j][r] = (r != j) ? -2 * v[j] * v[r] : 1 - 2 * v[j] * v[r];
#pragma unroll
    for (int k = 0; k < NR; ++k)
#pragma unroll
      for (int j = 0; j < NR - i; ++j) {
        double temp = 0;
#pragma unroll
        for (int r = 0; r < NR - i; ++r) temp += hh[j][r] * Ac[r + i][k];
        A[j + i][k] = temp;
      }
#pragma unroll
    for (int j = 0; j < NR - i; ++j) {
      double sum = 0;
#pragma unroll
      for (int r = 0; r < NR - i; ++r) sum += hh[j][r] * Bc[r + i];
      B[j + i] = sum;
    }
  }
  const double eps = DBL_EPSILON;
  double Bc[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) Bc[q] = B[q];
  int bottom = 0;
#pragma unroll
  for (int i = NR - 1; i >= 0; --i) {
    if (fabs(A[i][NR - 1]) > eps) {
      bottom = i;
      break;
    }
  }
#pragma unroll
  for (int i = NR - 1; i >= 0; --i) {
    if (i > bottom) continue;
    double sum = 0;
#pragma unroll
    for (int j = NR - 1; j > i; --j) sum += perm_get<NR>(B, jpvt[j]) * A[i][j];
    if (fabs(A[i][i]) > eps) {
      const double temp = 1 / A[i][i];
      perm_set<NR>(B, jpvt[i], (Bc[i] - sum) * temp);
    } else {
      perm_set<NR>(B, jpvt[i], 0.0);
    }
  }
}

// ------------------------------------------------------------ reductions
// max of non-negative doubles through their IEEE bit patterns
__device__ inline void atomic_max_nonneg(double *addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = smax(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace fsale
