// Gaussian-basis integrals for the host front-end (overlap, kinetic, nuclear
// attraction, dipole, electron repulsion) by the McMurchie-Davidson scheme
// (L. E. McMurchie, E. R. Davidson, J. Comput. Phys. 26, 218 (1978)):
// Hermite expansion coefficients E^{ij}_t of Gaussian products and Hermite
// Coulomb integrals R_{tuv} from the Boys function.  Shells up to l = 3;
// shells flagged "pure" are transformed to real solid harmonics (l = 2 only,
// which is what cc-pVDZ needs).  OpenMP over shell pairs / quartets.
//
// This replaces what libint2 provides to MPQC's AOFactory on the host
// (reference: src/mpqc/chemistry/qc/lcao/integrals/); it is not part of the
// GPU hot path.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <stdexcept>
#include <omp.h>

namespace {

constexpr int kMaxL = 3;
constexpr int kMaxLTot = 4 * kMaxL;
constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- Boys function
// F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt.  Tabulated on T = k*h (h = 1/32) for
// m <= kMaxLTot + 8 via the convergent series, evaluated by an 8-term Taylor
// expansion around the nearest grid point for the top m and downward
// recursion; asymptotic form with upward recursion beyond T = 40.
constexpr int kBoysM = kMaxLTot + 9;
constexpr double kBoysH = 1.0 / 32.0;
constexpr double kBoysTmax = 40.0;
constexpr int kBoysN = (int)(kBoysTmax / kBoysH) + 2;
std::vector<double> g_boys;  // [kBoysN][kBoysM]

double boys_series(int m, double T) {
  // F_m(T) = exp(-T) sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1))
  double term = 1.0 / (2 * m + 1), sum = term;
  for (int i = 1; i < 2000; ++i) {
    term *= 2.0 * T / (2 * m + 2 * i + 1);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  return std::exp(-T) * sum;
}

void init_boys() {
  if (!g_boys.empty()) return;
  std::vector<double> tab((size_t)kBoysN * kBoysM);
  for (int k = 0; k < kBoysN; ++k)
    for (int m = 0; m < kBoysM; ++m) tab[(size_t)k * kBoysM + m] = boys_series(m, k * kBoysH);
  g_boys.swap(tab);
}

// fills F[0..mmax]
inline void boys(int mmax, double T, double* F) {
  if (T < kBoysTmax) {
    const int k = (int)(T / kBoysH + 0.5);
    const double dT = k * kBoysH - T;  // F_m(T) = sum_j F_{m+j}(T_k) dT^j / j!
    const double* row = &g_boys[(size_t)k * kBoysM];
    double f = 0.0, fac = 1.0;
    for (int j = 0; j < 8; ++j) {
      f += row[mmax + j] * fac;
      fac *= dT / (j + 1);
    }
    F[mmax] = f;
    if (mmax > 0) {
      const double e = std::exp(-T);
      for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    }
  } else {
    const double e = std::exp(-T);
    F[0] = 0.5 * std::sqrt(kPi / T);
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
  }
}

// ---------------------------------------------------------------- shells
struct Shell {
  int l;
  bool pure;
  std::vector<double> a;  // exponents
  std::vector<double> c;  // contraction coefficients incl. primitive normalisation
  double R[3];
  int bf;  // first basis function (output basis)
};

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline int nfunc(const Shell& s) { return s.pure ? 2 * s.l + 1 : ncart(s.l); }

// Cartesian exponents in the order xx..., e.g. l=2: xx xy xz yy yz zz
void cart_list(int l, int (*out)[3]) {
  int n = 0;
  for (int i = l; i >= 0; --i)
    for (int j = l - i; j >= 0; --j) {
      out[n][0] = i; out[n][1] = j; out[n][2] = l - i - j;
      ++n;
    }
}

double dfact(int n) {  // (n)!! with (-1)!! = 1
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// Cartesian -> real solid harmonics for normalised Cartesian components.
// Rows: spherical functions; columns: Cartesian components (cart_list order).
void sph_matrix(int l, std::vector<double>& M) {
  if (l == 2) {
    const double s3 = std::sqrt(3.0);
    // xx xy xz yy yz zz
    M = {0, 1, 0, 0, 0, 0,                 // xy
         0, 0, 0, 0, 1, 0,                 // yz
         -0.5, 0, 0, -0.5, 0, 1,           // z^2
         0, 0, 1, 0, 0, 0,                 // xz
         s3 / 2, 0, 0, -s3 / 2, 0, 0};     // x^2 - y^2
  } else {
    throw std::runtime_error("pure shells supported for l = 2 only");
  }
}

// Hermite coefficients E^{ij}_t for one Cartesian direction;
// E[i][j][t], i <= la, j <= lb, t <= i + j.
struct E1 {
  double v[kMaxL + 1][kMaxL + 3][2 * kMaxL + 4];
};

inline void hermite_E(int la, int lb, double p, double XPA, double XPB, double E00, E1& E) {
  const double oo2p = 0.5 / p;
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j)
      for (int t = 0; t <= i + j + 1; ++t) E.v[i][j][t] = 0.0;
  E.v[0][0][0] = E00;
  for (int i = 0; i <= la; ++i) {
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      if (j == 0) {  // from (i-1, 0)
        for (int t = 0; t <= i + j; ++t) {
          double r = XPA * E.v[i - 1][j][t];
          if (t > 0) r += oo2p * E.v[i - 1][j][t - 1];
          if (t + 1 <= i - 1 + j) r += (t + 1) * E.v[i - 1][j][t + 1];
          E.v[i][j][t] = r;
        }
      } else {  // from (i, j-1)
        for (int t = 0; t <= i + j; ++t) {
          double r = XPB * E.v[i][j - 1][t];
          if (t > 0) r += oo2p * E.v[i][j - 1][t - 1];
          if (t + 1 <= i + j - 1) r += (t + 1) * E.v[i][j - 1][t + 1];
          E.v[i][j][t] = r;
        }
      }
    }
  }
}

// Hermite Coulomb integrals R_{tuv}(alpha, X, Y, Z) for t+u+v <= L.
struct RTab {
  int L;
  double v[kMaxLTot + 1][kMaxLTot + 1][kMaxLTot + 1];
};

inline void hermite_R(int L, double alpha, double X, double Y, double Z, RTab& out) {
  double F[kMaxLTot + 1];
  boys(L, alpha * (X * X + Y * Y + Z * Z), F);
  // Rn[n][t][u][v] with t+u+v <= L - n; compute top-down in n
  static thread_local double Rn[kMaxLTot + 1][kMaxLTot + 1][kMaxLTot + 1][kMaxLTot + 1];
  double m2a = 1.0;
  for (int n = 0; n <= L; ++n) {
    Rn[n][0][0][0] = m2a * F[n];
    m2a *= -2.0 * alpha;
  }
  for (int n = L - 1; n >= 0; --n) {
    const int top = L - n;
    for (int t = 0; t <= top; ++t)
      for (int u = 0; u <= top - t; ++u)
        for (int v = 0; v <= top - t - u; ++v) {
          if (t == 0 && u == 0 && v == 0) continue;
          double r;
          if (t > 0) {
            r = X * Rn[n + 1][t - 1][u][v];
            if (t > 1) r += (t - 1) * Rn[n + 1][t - 2][u][v];
          } else if (u > 0) {
            r = Y * Rn[n + 1][t][u - 1][v];
            if (u > 1) r += (u - 1) * Rn[n + 1][t][u - 2][v];
          } else {
            r = Z * Rn[n + 1][t][u][v - 1];
            if (v > 1) r += (v - 1) * Rn[n + 1][t][u][v - 2];
          }
          Rn[n][t][u][v] = r;
        }
  }
  out.L = L;
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) out.v[t][u][v] = Rn[0][t][u][v];
}

// Primitive-pair data of a shell pair.
struct PrimPair {
  double p, P[3];
  E1 Ex, Ey, Ez;  // contraction coefficients folded into Ex[0][0][0]
};

struct ShellPair {
  int a, b;  // shell indices
  std::vector<PrimPair> pp;
};

void make_pair(const Shell& A, const Shell& B, int ia, int ib, ShellPair& sp) {
  sp.a = ia; sp.b = ib;
  sp.pp.clear();
  const double AB[3] = {A.R[0] - B.R[0], A.R[1] - B.R[1], A.R[2] - B.R[2]};
  const double AB2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  for (size_t i = 0; i < A.a.size(); ++i)
    for (size_t j = 0; j < B.a.size(); ++j) {
      const double a = A.a[i], b = B.a[j], p = a + b, mu = a * b / p;
      const double K = std::exp(-mu * AB2);
      const double pref = A.c[i] * B.c[j] * K;
      if (std::fabs(pref) < 1e-22) continue;
      PrimPair q;
      q.p = p;
      for (int d = 0; d < 3; ++d) q.P[d] = (a * A.R[d] + b * B.R[d]) / p;
      hermite_E(A.l, B.l, p, q.P[0] - A.R[0], q.P[0] - B.R[0], pref, q.Ex);
      hermite_E(A.l, B.l, p, q.P[1] - A.R[1], q.P[1] - B.R[1], 1.0, q.Ey);
      hermite_E(A.l, B.l, p, q.P[2] - A.R[2], q.P[2] - B.R[2], 1.0, q.Ez);
      sp.pp.push_back(q);
    }
}

// transform a Cartesian block (na_c x nb_c row-major) of a shell pair into the
// output basis and add into out (ld = nbf)
template <typename F>
void emit_pair(const Shell& A, const Shell& B, const double* cart, F&& store) {
  const int nca = ncart(A.l), ncb = ncart(B.l);
  std::vector<double> MA, MB;
  if (A.pure) sph_matrix(A.l, MA);
  if (B.pure) sph_matrix(B.l, MB);
  const int na = nfunc(A), nb = nfunc(B);
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      double s = 0.0;
      for (int x = 0; x < nca; ++x) {
        const double ca = A.pure ? MA[i * nca + x] : (x == i ? 1.0 : 0.0);
        if (ca == 0.0) continue;
        for (int y = 0; y < ncb; ++y) {
          const double cb = B.pure ? MB[j * ncb + y] : (y == j ? 1.0 : 0.0);
          if (cb == 0.0) continue;
          s += ca * cb * cart[x * ncb + y];
        }
      }
      store(A.bf + i, B.bf + j, s);
    }
}

std::vector<Shell> make_shells(int nshell, const int* l, const int* pure, const int* nprim,
                               const double* exps, const double* coefs, const double* centers,
                               const int* bf) {
  std::vector<Shell> sh(nshell);
  int off = 0;
  for (int s = 0; s < nshell; ++s) {
    if (l[s] < 0 || l[s] > kMaxL) throw std::runtime_error("angular momentum out of range");
    Shell& x = sh[s];
    x.l = l[s];
    x.pure = pure[s] != 0 && l[s] >= 2;
    x.a.assign(exps + off, exps + off + nprim[s]);
    // primitive normalisation of the x^l component times the contraction
    // coefficient, then normalise the contraction
    x.c.resize(nprim[s]);
    const int L = x.l;
    for (int k = 0; k < nprim[s]; ++k) {
      const double a = x.a[k];
      const double N = std::pow(2.0 * a / kPi, 0.75) * std::pow(4.0 * a, 0.5 * L) /
                       std::sqrt(dfact(2 * L - 1));
      x.c[k] = coefs[off + k] * N;
    }
    double norm = 0.0;
    for (int i = 0; i < nprim[s]; ++i)
      for (int j = 0; j < nprim[s]; ++j) {
        const double p = x.a[i] + x.a[j];
        norm += x.c[i] * x.c[j] * std::pow(kPi / p, 1.5) * dfact(2 * L - 1) / std::pow(2.0 * p, L);
      }
    for (auto& c : x.c) c /= std::sqrt(norm);
    for (int d = 0; d < 3; ++d) x.R[d] = centers[3 * s + d];
    x.bf = bf[s];
    off += nprim[s];
  }
  return sh;
}

// component normalisation relative to the x^l component:
// sqrt((2l-1)!! / ((2i-1)!!(2j-1)!!(2k-1)!!))
double comp_norm(int i, int j, int k) {
  const int l = i + j + k;
  return std::sqrt(dfact(2 * l - 1) / (dfact(2 * i - 1) * dfact(2 * j - 1) * dfact(2 * k - 1)));
}


// Cartesian (ab|cd) of one shell quartet (bra pair a b, ket pair c d), component-normalised.
void quartet_cart(const std::vector<Shell>& sh, const ShellPair& bra, const ShellPair& ket,
                std::vector<double>& cart) {
  const Shell& A = sh[bra.a];
  const Shell& B = sh[bra.b];
  const Shell& C = sh[ket.a];
  const Shell& D = sh[ket.b];
  int ca[10][3], cb[10][3], cc[10][3], cd[10][3];
  cart_list(A.l, ca); cart_list(B.l, cb); cart_list(C.l, cc); cart_list(D.l, cd);
  const int nca = ncart(A.l), ncb = ncart(B.l), ncc = ncart(C.l), ncd = ncart(D.l);
  const int Lab = A.l + B.l, Lcd = C.l + D.l, L = Lab + Lcd;
  const int nket = ncc * ncd;
  // W[t][u][v][ket cart pair]
  const int dimH = Lab + 1;
  std::vector<double> W((size_t)dimH * dimH * dimH * nket, 0.0);
  cart.assign((size_t)nca * ncb * nket, 0.0);
  RTab R;
  for (const PrimPair& P : bra.pp) {
    std::fill(W.begin(), W.end(), 0.0);
    for (const PrimPair& Q : ket.pp) {
      const double p = P.p, q = Q.p;
      const double alpha = p * q / (p + q);
      hermite_R(L, alpha, P.P[0] - Q.P[0], P.P[1] - Q.P[1], P.P[2] - Q.P[2], R);
      const double pref = 2.0 * std::pow(kPi, 2.5) / (p * q * std::sqrt(p + q));
      for (int z = 0; z < ncc; ++z)
        for (int w = 0; w < ncd; ++w) {
          const int i2 = cc[z][0], j2 = cc[z][1], k2 = cc[z][2];
          const int i3 = cd[w][0], j3 = cd[w][1], k3 = cd[w][2];
          const int kp = z * ncd + w;
          for (int t = 0; t <= Lab; ++t)
            for (int u = 0; u <= Lab - t; ++u)
              for (int v = 0; v <= Lab - t - u; ++v) {
                double s = 0.0;
                for (int tau = 0; tau <= i2 + i3; ++tau) {
                  const double ex = Q.Ex.v[i2][i3][tau];
                  if (ex == 0.0) continue;
                  for (int nu = 0; nu <= j2 + j3; ++nu) {
                    const double exy = ex * Q.Ey.v[j2][j3][nu];
                    if (exy == 0.0) continue;
                    for (int phi = 0; phi <= k2 + k3; ++phi) {
                      const double e = exy * Q.Ez.v[k2][k3][phi];
                      const double r = R.v[t + tau][u + nu][v + phi];
                      s += ((tau + nu + phi) & 1) ? -e * r : e * r;
                    }
                  }
                }
                W[(((size_t)t * dimH + u) * dimH + v) * nket + kp] += pref * s;
              }
        }
    }
    for (int x = 0; x < nca; ++x)
      for (int y = 0; y < ncb; ++y) {
        const int i0 = ca[x][0], j0 = ca[x][1], k0 = ca[x][2];
        const int i1 = cb[y][0], j1 = cb[y][1], k1 = cb[y][2];
        double* dst = &cart[((size_t)x * ncb + y) * nket];
        for (int t = 0; t <= i0 + i1; ++t) {
          const double ex = P.Ex.v[i0][i1][t];
          if (ex == 0.0) continue;
          for (int u = 0; u <= j0 + j1; ++u) {
            const double exy = ex * P.Ey.v[j0][j1][u];
            if (exy == 0.0) continue;
            for (int v = 0; v <= k0 + k1; ++v) {
              const double e = exy * P.Ez.v[k0][k1][v];
              if (e == 0.0) continue;
              const double* src = &W[(((size_t)t * dimH + u) * dimH + v) * nket];
              for (int kp = 0; kp < nket; ++kp) dst[kp] += e * src[kp];
            }
          }
        }
      }
  }
  // Cartesian component normalisation
  for (int x = 0; x < nca; ++x)
    for (int y = 0; y < ncb; ++y)
      for (int z = 0; z < ncc; ++z)
        for (int w = 0; w < ncd; ++w)
          cart[(((size_t)x * ncb + y) * ncc + z) * ncd + w] *=
              comp_norm(ca[x][0], ca[x][1], ca[x][2]) * comp_norm(cb[y][0], cb[y][1], cb[y][2]) *
              comp_norm(cc[z][0], cc[z][1], cc[z][2]) * comp_norm(cd[w][0], cd[w][1], cd[w][2]);
}

// transform a Cartesian quartet to the output basis: out4[na][nb][nc][nd]
void quartet_to_output(const Shell* const S4[4], const std::vector<double>& cart, std::vector<double>& res) {
  int nc[4], nf[4];
  std::vector<double> M[4];
  for (int k = 0; k < 4; ++k) {
    nc[k] = ncart(S4[k]->l);
    nf[k] = nfunc(*S4[k]);
    if (S4[k]->pure) sph_matrix(S4[k]->l, M[k]);
  }
  std::vector<double> cur(cart), nxt;
  int dims[4] = {nc[0], nc[1], nc[2], nc[3]};
  for (int k = 0; k < 4; ++k) {
    if (!S4[k]->pure) continue;
    int nd[4] = {dims[0], dims[1], dims[2], dims[3]};
    nd[k] = nf[k];
    nxt.assign((size_t)nd[0] * nd[1] * nd[2] * nd[3], 0.0);
    for (int a = 0; a < nd[0]; ++a)
      for (int b = 0; b < nd[1]; ++b)
        for (int c = 0; c < nd[2]; ++c)
          for (int d = 0; d < nd[3]; ++d) {
            int idx[4] = {a, b, c, d};
            const int f = idx[k];
            double s = 0.0;
            for (int x = 0; x < nc[k]; ++x) {
              const double m = M[k][f * nc[k] + x];
              if (m == 0.0) continue;
              int j[4] = {a, b, c, d};
              j[k] = x;
              s += m * cur[(((size_t)j[0] * dims[1] + j[1]) * dims[2] + j[2]) * dims[3] + j[3]];
            }
            nxt[(((size_t)a * nd[1] + b) * nd[2] + c) * nd[3] + d] = s;
          }
    cur.swap(nxt);
    for (int q = 0; q < 4; ++q) dims[q] = nd[q];
  }
  res.swap(cur);
}
}  // namespace

extern "C" {

// One-electron integrals in the output basis (nbf x nbf row-major): overlap S,
// kinetic T, nuclear attraction V (sum over point charges Z at xyz), and the
// dipole components Dx, Dy, Dz about `origin` (any of the outputs may be NULL).
int chem_one_electron(int nshell, const int* l, const int* pure, const int* nprim, const double* exps,
                      const double* coefs, const double* centers, const int* bf, int nbf, int natom,
                      const double* Z, const double* xyz, const double* origin, double* S, double* T,
                      double* V, double* Dx, double* Dy, double* Dz) {
  try {
    init_boys();
    auto sh = make_shells(nshell, l, pure, nprim, exps, coefs, centers, bf);
    double* outs[6] = {S, T, V, Dx, Dy, Dz};
    for (double* o : outs)
      if (o) std::memset(o, 0, sizeof(double) * (size_t)nbf * nbf);
    std::vector<std::pair<int, int>> pairs;
    for (int a = 0; a < nshell; ++a)
      for (int b = 0; b <= a; ++b) pairs.push_back({a, b});
#pragma omp parallel for schedule(dynamic)
    for (long k = 0; k < (long)pairs.size(); ++k) {
      const int ia = pairs[k].first, ib = pairs[k].second;
      const Shell& A = sh[ia];
      const Shell& B = sh[ib];
      ShellPair sp;
      make_pair(A, B, ia, ib, sp);
      int ca[10][3], cb[10][3];
      cart_list(A.l, ca);
      cart_list(B.l, cb);
      const int nca = ncart(A.l), ncb = ncart(B.l);
      std::vector<double> blk[6];
      for (auto& v : blk) v.assign((size_t)nca * ncb, 0.0);
      for (const PrimPair& q : sp.pp) {
        const double p = q.p;
        const double b = 0;  // unused
        (void)b;
        const double sq = std::sqrt(kPi / p);
        for (int x = 0; x < nca; ++x)
          for (int y = 0; y < ncb; ++y) {
            const int i1 = ca[x][0], j1 = ca[x][1], k1 = ca[x][2];
            const int i2 = cb[y][0], j2 = cb[y][1], k2 = cb[y][2];
            const double nrm = comp_norm(i1, j1, k1) * comp_norm(i2, j2, k2);
            const double sx = q.Ex.v[i1][i2][0] * sq, sy = q.Ey.v[j1][j2][0] * sq,
                         sz = q.Ez.v[k1][k2][0] * sq;
            if (S) blk[0][x * ncb + y] += nrm * sx * sy * sz;
            if (Dx || Dy || Dz) {
              const double mx = (q.Ex.v[i1][i2][1] + (q.P[0] - origin[0]) * q.Ex.v[i1][i2][0]) * sq;
              const double my = (q.Ey.v[j1][j2][1] + (q.P[1] - origin[1]) * q.Ey.v[j1][j2][0]) * sq;
              const double mz = (q.Ez.v[k1][k2][1] + (q.P[2] - origin[2]) * q.Ez.v[k1][k2][0]) * sq;
              blk[3][x * ncb + y] += nrm * mx * sy * sz;
              blk[4][x * ncb + y] += nrm * sx * my * sz;
              blk[5][x * ncb + y] += nrm * sx * sy * mz;
            }
          }
        if (T) {
          // kinetic energy via overlaps with raised/lowered B exponents:
          // T_1d(i,j) = -2 b^2 S(i,j+2) + b (2j+1) S(i,j) - j(j-1)/2 S(i,j-2)
          // (needs E up to lb+2: recompute locally)
          // handled below with a second pass per primitive pair
        }
      }
      if (T) {
        // recompute primitive pairs with lb + 2 to obtain the kinetic energy
        for (size_t ip = 0; ip < A.a.size(); ++ip)
          for (size_t jp = 0; jp < B.a.size(); ++jp) {
            const double a = A.a[ip], bb = B.a[jp], p = a + bb, mu = a * bb / p;
            double P[3], AB2 = 0.0;
            for (int d = 0; d < 3; ++d) {
              P[d] = (a * A.R[d] + bb * B.R[d]) / p;
              AB2 += (A.R[d] - B.R[d]) * (A.R[d] - B.R[d]);
            }
            const double pref = A.c[ip] * B.c[jp] * std::exp(-mu * AB2);
            E1 E[3];
            for (int d = 0; d < 3; ++d) {
              // E with lb + 2 (E1 sized for kMaxL + 2 in j)
              const int la = A.l, lb = B.l + 2;
              const double oo2p = 0.5 / p, XPA = P[d] - A.R[d], XPB = P[d] - B.R[d];
              for (int i = 0; i <= la; ++i)
                for (int j = 0; j <= lb; ++j)
                  for (int t = 0; t < 2 * kMaxL + 4; ++t) E[d].v[i][j][t] = 0.0;
              E[d].v[0][0][0] = 1.0;
              for (int i = 0; i <= la; ++i)
                for (int j = 0; j <= lb; ++j) {
                  if (i == 0 && j == 0) continue;
                  for (int t = 0; t <= i + j; ++t) {
                    double r;
                    if (j == 0) {
                      r = XPA * E[d].v[i - 1][j][t];
                      if (t > 0) r += oo2p * E[d].v[i - 1][j][t - 1];
                      if (t + 1 <= i - 1 + j) r += (t + 1) * E[d].v[i - 1][j][t + 1];
                    } else {
                      r = XPB * E[d].v[i][j - 1][t];
                      if (t > 0) r += oo2p * E[d].v[i][j - 1][t - 1];
                      if (t + 1 <= i + j - 1) r += (t + 1) * E[d].v[i][j - 1][t + 1];
                    }
                    E[d].v[i][j][t] = r;
                  }
                }
            }
            const double sq = std::sqrt(kPi / p);
            auto S1 = [&](int d, int i, int j) { return j < 0 ? 0.0 : E[d].v[i][j][0] * sq; };
            auto T1 = [&](int d, int i, int j) {
              return -2.0 * bb * bb * S1(d, i, j + 2) + bb * (2 * j + 1) * S1(d, i, j) -
                     0.5 * j * (j - 1) * S1(d, i, j - 2);
            };
            for (int x = 0; x < nca; ++x)
              for (int y = 0; y < ncb; ++y) {
                const int i1 = ca[x][0], j1 = ca[x][1], k1 = ca[x][2];
                const int i2 = cb[y][0], j2 = cb[y][1], k2 = cb[y][2];
                const double nrm = comp_norm(i1, j1, k1) * comp_norm(i2, j2, k2);
                const double t = T1(0, i1, i2) * S1(1, j1, j2) * S1(2, k1, k2) +
                                 S1(0, i1, i2) * T1(1, j1, j2) * S1(2, k1, k2) +
                                 S1(0, i1, i2) * S1(1, j1, j2) * T1(2, k1, k2);
                blk[1][x * ncb + y] += pref * nrm * t;
              }
          }
      }
      if (V) {
        for (const PrimPair& q : sp.pp) {
          const int L = A.l + B.l;
          for (int c = 0; c < natom; ++c) {
            RTab R;
            hermite_R(L, q.p, q.P[0] - xyz[3 * c], q.P[1] - xyz[3 * c + 1], q.P[2] - xyz[3 * c + 2], R);
            const double pref = -Z[c] * 2.0 * kPi / q.p;
            for (int x = 0; x < nca; ++x)
              for (int y = 0; y < ncb; ++y) {
                const int i1 = ca[x][0], j1 = ca[x][1], k1 = ca[x][2];
                const int i2 = cb[y][0], j2 = cb[y][1], k2 = cb[y][2];
                const double nrm = comp_norm(i1, j1, k1) * comp_norm(i2, j2, k2);
                double s = 0.0;
                for (int t = 0; t <= i1 + i2; ++t)
                  for (int u = 0; u <= j1 + j2; ++u)
                    for (int v = 0; v <= k1 + k2; ++v)
                      s += q.Ex.v[i1][i2][t] * q.Ey.v[j1][j2][u] * q.Ez.v[k1][k2][v] * R.v[t][u][v];
                blk[2][x * ncb + y] += pref * nrm * s;
              }
          }
        }
      }
      for (int w = 0; w < 6; ++w) {
        double* o = outs[w];
        if (!o) continue;
        emit_pair(A, B, blk[w].data(), [&](int i, int j, double val) {
          o[(size_t)i * nbf + j] = val;
          o[(size_t)j * nbf + i] = val;
        });
      }
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

// Electron repulsion integrals (ij|kl) (chemist notation) in the output basis,
// written densely to out[nbf^4] (row-major i,j,k,l).  Shell quartets whose
// Schwarz bound is below `screen` are skipped (their integrals set to 0).
int chem_eri(int nshell, const int* l, const int* pure, const int* nprim, const double* exps,
             const double* coefs, const double* centers, const int* bf, int nbf, double screen,
             double* out) {
  try {
    init_boys();
    auto sh = make_shells(nshell, l, pure, nprim, exps, coefs, centers, bf);
    const size_t n = (size_t)nbf;
    std::memset(out, 0, sizeof(double) * n * n * n * n);
    // shell pairs (a >= b)
    std::vector<ShellPair> pairs;
    for (int a = 0; a < nshell; ++a)
      for (int b = 0; b <= a; ++b) {
        ShellPair sp;
        make_pair(sh[a], sh[b], a, b, sp);
        pairs.push_back(std::move(sp));
      }
    const long np = (long)pairs.size();
    std::vector<double> schwarz(np, 0.0);

    auto quartet = [&](const ShellPair& bra, const ShellPair& ket, std::vector<double>& cart) {
      quartet_cart(sh, bra, ket, cart);
    };

    auto to_output = [](const Shell* const S4[4], const std::vector<double>& cart, std::vector<double>& res) {
      quartet_to_output(S4, cart, res);
    };

    // Schwarz factors sqrt(max |(ab|ab)|)
#pragma omp parallel for schedule(dynamic)
    for (long k = 0; k < np; ++k) {
      std::vector<double> cart, res;
      quartet(pairs[k], pairs[k], cart);
      const Shell* S4[4] = {&sh[pairs[k].a], &sh[pairs[k].b], &sh[pairs[k].a], &sh[pairs[k].b]};
      to_output(S4, cart, res);
      double m = 0.0;
      for (double x : res) m = std::max(m, std::fabs(x));
      schwarz[k] = std::sqrt(m);
    }

    // unique quartets: bra pair index >= ket pair index
#pragma omp parallel for schedule(dynamic, 4)
    for (long kb = 0; kb < np; ++kb) {
      std::vector<double> cart, res;
      for (long kk = 0; kk <= kb; ++kk) {
        if (schwarz[kb] * schwarz[kk] < screen) continue;
        const ShellPair& bra = pairs[kb];
        const ShellPair& ket = pairs[kk];
        quartet(bra, ket, cart);
        const Shell* S4[4] = {&sh[bra.a], &sh[bra.b], &sh[ket.a], &sh[ket.b]};
        to_output(S4, cart, res);
        const int na = nfunc(*S4[0]), nb = nfunc(*S4[1]), nc = nfunc(*S4[2]), nd = nfunc(*S4[3]);
        for (int a = 0; a < na; ++a)
          for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c)
              for (int d = 0; d < nd; ++d) {
                const double val = res[(((size_t)a * nb + b) * nc + c) * nd + d];
                const size_t i = S4[0]->bf + a, j = S4[1]->bf + b, k = S4[2]->bf + c,
                             l2 = S4[3]->bf + d;
                // 8-fold permutational symmetry; duplicates write the same value
                out[((i * n + j) * n + k) * n + l2] = val;
                out[((j * n + i) * n + k) * n + l2] = val;
                out[((i * n + j) * n + l2) * n + k] = val;
                out[((j * n + i) * n + l2) * n + k] = val;
                out[((k * n + l2) * n + i) * n + j] = val;
                out[((l2 * n + k) * n + i) * n + j] = val;
                out[((k * n + l2) * n + j) * n + i] = val;
                out[((l2 * n + k) * n + j) * n + i] = val;
              }
      }
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

}  // extern "C"

extern "C" {
// Density-fitting Coulomb integrals: three-centre (P|mu nu) over auxiliary
// shells P and orbital shells mu, nu, written to out[naux][nbf][nbf], and
// two-centre (P|Q) written to out[naux][naux].  The missing fourth function is
// the constant 1 (an s primitive of exponent 0 with coefficient 1), so each
// block is the quartet (P 1|mu nu) or (P 1|Q 1) of the four-centre code.
int chem_eri3(int nshell, const int* l, const int* pure, const int* nprim, const double* exps,
              const double* coefs, const double* centers, const int* bf, int nbf, int nashell,
              const int* al, const int* apure, const int* anprim, const double* aexps,
              const double* acoefs, const double* acenters, const int* abf, int naux, double* out) {
  try {
    init_boys();
    auto sh = make_shells(nshell, l, pure, nprim, exps, coefs, centers, bf);
    auto ash = make_shells(nashell, al, apure, anprim, aexps, acoefs, acenters, abf);
    const int a0 = (int)sh.size();
    sh.insert(sh.end(), ash.begin(), ash.end());
    Shell unit;
    unit.l = 0; unit.pure = false; unit.a = {0.0}; unit.c = {1.0};
    unit.R[0] = unit.R[1] = unit.R[2] = 0.0; unit.bf = 0;
    const int iu = (int)sh.size();
    sh.push_back(unit);
    std::vector<ShellPair> aux(nashell), orb;
    for (int P = 0; P < nashell; ++P) {
      sh[iu].R[0] = sh[a0 + P].R[0]; sh[iu].R[1] = sh[a0 + P].R[1]; sh[iu].R[2] = sh[a0 + P].R[2];
      make_pair(sh[a0 + P], sh[iu], a0 + P, iu, aux[P]);
    }
    for (int a = 0; a < nshell; ++a)
      for (int b = 0; b <= a; ++b) {
        ShellPair sp;
        make_pair(sh[a], sh[b], a, b, sp);
        orb.push_back(std::move(sp));
      }
    const size_t n = (size_t)nbf;
    const long nwork = (long)nashell * (long)orb.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (long w = 0; w < nwork; ++w) {
      const int P = (int)(w / (long)orb.size());
      const ShellPair& ket = orb[w % (long)orb.size()];
      std::vector<double> cart, res;
      quartet_cart(sh, aux[P], ket, cart);
      const Shell* S4[4] = {&sh[a0 + P], &sh[iu], &sh[ket.a], &sh[ket.b]};
      quartet_to_output(S4, cart, res);
      const int np = nfunc(*S4[0]), nc = nfunc(*S4[2]), nd = nfunc(*S4[3]);
      for (int x = 0; x < np; ++x)
        for (int c = 0; c < nc; ++c)
          for (int d = 0; d < nd; ++d) {
            const double val = res[((size_t)x * nc + c) * nd + d];
            const size_t q = S4[0]->bf + x, i = S4[2]->bf + c, j = S4[3]->bf + d;
            out[(q * n + i) * n + j] = val;
            out[(q * n + j) * n + i] = val;
          }
    }
    (void)naux;
    return 0;
  } catch (...) {
    return 1;
  }
}

int chem_eri2(int nashell, const int* al, const int* apure, const int* anprim, const double* aexps,
              const double* acoefs, const double* acenters, const int* abf, int naux, double* out) {
  try {
    init_boys();
    auto sh = make_shells(nashell, al, apure, anprim, aexps, acoefs, acenters, abf);
    Shell unit;
    unit.l = 0; unit.pure = false; unit.a = {0.0}; unit.c = {1.0}; unit.bf = 0;
    const int iu = (int)sh.size();
    sh.push_back(unit);
    std::vector<ShellPair> aux(nashell);
    for (int P = 0; P < nashell; ++P) {
      for (int d = 0; d < 3; ++d) sh[iu].R[d] = sh[P].R[d];
      make_pair(sh[P], sh[iu], P, iu, aux[P]);
    }
    const size_t n = (size_t)naux;
#pragma omp parallel for schedule(dynamic)
    for (int P = 0; P < nashell; ++P) {
      std::vector<double> cart, res;
      for (int Q = 0; Q <= P; ++Q) {
        quartet_cart(sh, aux[P], aux[Q], cart);
        const Shell* S4[4] = {&sh[P], &sh[iu], &sh[Q], &sh[iu]};
        quartet_to_output(S4, cart, res);
        const int np = nfunc(sh[P]), nq = nfunc(sh[Q]);
        for (int x = 0; x < np; ++x)
          for (int y = 0; y < nq; ++y) {
            const double val = res[(size_t)x * nq + y];
            const size_t i = sh[P].bf + x, j = sh[Q].bf + y;
            out[i * n + j] = val;
            out[j * n + i] = val;
          }
      }
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

}  // extern "C"

extern "C" {
// Coulomb and exchange matrices from dense (ij|kl) and a symmetric density:
// J_ij = sum_kl (ij|kl) D_kl,  K_ij = sum_kl (ik|jl) D_kl.  One pass over the
// integrals, OpenMP over i.
int chem_jk(int n, const double* eri, const double* D, double* J, double* K) {
  const size_t N = (size_t)n;
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      const double* e = eri + (((size_t)i * N + j) * N) * N;
      double s = 0.0;
      for (size_t kl = 0; kl < N * N; ++kl) s += e[kl] * D[kl];
      J[(size_t)i * N + j] = s;
    }
    std::vector<double> row(N, 0.0);
    for (int k = 0; k < n; ++k) {
      const double* Dk = D + (size_t)k * N;
      for (int j = 0; j < n; ++j) {
        const double* e = eri + (((size_t)i * N + k) * N + j) * N;
        double s = 0.0;
        for (int l = 0; l < n; ++l) s += e[l] * Dk[l];
        row[j] += s;
      }
    }
    for (int j = 0; j < n; ++j) K[(size_t)i * N + j] = row[j];
  }
  return 0;
}
}  // extern "C"
