// Host tables of the plane-wave FMM (pw.hpp): plane-wave quadrature, the
// rotation of expansions into the six wave directions, and the dense
// per-level operators M2M, L2L, M2X and X2L, each built by projecting the
// exact translated function onto the basis on a sphere (Gauss-Legendre in
// cos(theta), trapezoid in phi).  Follows the roles of the reference's
// Level_Dependent_FMM_Globals (src/fmm/fmm_globals.h:372-493: quadrature,
// ympshiftcoef, ylcshftcoef, yrlscini, ymkexps, ymkfexp); the operators are
// formed as dense matrices for batched GEMMs instead of rotate-shift-rotate
// and Fourier sequences.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <map>
#include <mutex>
#include <random>
#include <stdexcept>
#include <tuple>

#include "pw.hpp"

namespace beepdev {

namespace {

constexpr double PI = 3.14159265358979323846;
using cd = std::complex<double>;

// Laplace plane-wave quadrature (generalized Gaussian rules of Yarvin and
// Rokhlin for  1/r = int_0^inf e^{-t z} J_0(t rho) dt,  1 <= z <= 4,
// rho <= 4 sqrt2; 9 nodes: about 3 digits, 18 nodes: about 6), with the
// angles each node needs in the J_0 integral.  Published constants, the same
// as src/fmm/fmm_globals.h:532-577 (nodes, weights) and :580-705 (angle counts).
const double YR9_T[9] = {.099273996739714473469540223504736787, .47725674637049431137114652301534079,
                         1.0553366138218296388373573790886439, 1.7675934335400844688024335482623428,
                         2.5734262935147067530294862081063911, 3.4482433920158257478760788217186928,
                         4.3768098355472631055818055756390095, 5.3489575720546005399569367000367492,
                         6.3576578531337464283978988532908261};
const double YR9_W[9] = {.24776441819008371281185532097879332, .49188566500464336872511239562300034,
                         .65378749137677805158830324216978624, .76433038408784093054038066838984378,
                         .84376180565628111640563702167128213, .90445883985098263213586733400006779,
                         .9537861313683345665381807521043811, .99670261613218547047665651916759089,
                         1.0429422730252668749528766056755558};
const int YR9_NPHYS[9] = {4, 8, 12, 16, 20, 20, 24, 8, 2};
const double YR18_T[18] = {
    .052788527661177607475107009804560221, .26949859838931256028615734976483509,
    .6322035317468939208396250251098536,   1.1130756427760852833586113774799742,
    1.6893949614021379623807206371566281,  2.3437620046953044905535534780938178,
    3.0626998290780611533534738555317745,  3.8356294126529686394633245072327554,
    4.6542473432156272750148673367220908,  5.5120938659358147404532246582675725,
    6.4042126837727888499784967279992998,  7.3268800190617540124549122992902994,
    8.2774009925823861522076185792684555,  9.2539718060248947750778825138695538,
    10.255602723746401139237605093512684,  11.282088297877740146191172243561596,
    12.334067909676926788620221486780792,  13.414920240172401477707353478763252};
const double YR18_W[18] = {
    .13438265914335215112096477696468355, .29457752727395436487256574764614925,
    .42607819361148618897416895379137713, .53189220776549905878027857397682965,
    .61787306245538586857435348065337166, .68863156078905074508611505734734237,
    .74749099381426187260757387775811367, .79699192718599998208617307682288811,
    .83917454386997591964103548889397644, .8757009228374531550898041132313665,
    .90792943590067498593754180546966381, .93698393742461816291466902839601971,
    .96382546688788062194674921556725167, .98932985769673820186653756536543369,
    1.0143828459791703888726033255807124, 1.0400365437416452252250564924906939,
    1.0681548926956736522697610780596733, 1.1090758097553685690428437737864442};
const int YR18_NPHYS[18] = {6, 8, 12, 16, 20, 26, 30, 34, 38, 44, 48, 52, 56, 60, 60, 52, 4, 2};

// psi_j(v) for all j <= p (real basis, see pw.hpp), with gradient if g3.
void basis(int p, double beta, const double v[3], double* psi, double (*g3)[3] = nullptr) {
    double g[PW_MAXP + 2];
    const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    pw_g<double, PW_MAXP>(beta * beta * r2, g);
    double sr[PW_MAXP + 1], si[PW_MAXP + 1], gr[PW_MAXP + 1][3], gi[PW_MAXP + 1][3];
    const double rt2 = std::sqrt(2.0);
    for (int m = 0; m <= p; ++m) {
        pw_column<double, PW_MAXP>(m, v[0], v[1], v[2], sr, si, gr, gi);
        for (int n = m; n <= p; ++n) {
            // d g_n / d v = beta^2 g_{n+1} / (2n+3) v
            const double dg = beta * beta * g[n + 1] / double(2 * n + 3);
            if (m == 0) {
                psi[pw_idx(n, 0, 0)] = g[n] * sr[n];
                if (g3)
                    for (int d = 0; d < 3; ++d) g3[pw_idx(n, 0, 0)][d] = g[n] * gr[n][d] + dg * sr[n] * v[d];
            } else {
                psi[pw_idx(n, m, 0)] = rt2 * g[n] * sr[n];
                psi[pw_idx(n, m, 1)] = rt2 * g[n] * si[n];
                if (g3)
                    for (int d = 0; d < 3; ++d) {
                        g3[pw_idx(n, m, 0)][d] = rt2 * (g[n] * gr[n][d] + dg * sr[n] * v[d]);
                        g3[pw_idx(n, m, 1)][d] = rt2 * (g[n] * gi[n][d] + dg * si[n] * v[d]);
                    }
            }
        }
    }
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int k = 1; k <= n; ++k) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double dz = p1 / pp;
            z -= dz;
            if (std::abs(dz) < 1e-16) break;
        }
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

// Sphere quadrature of radius rho: points and weights (dOmega) with the
// basis psi_j of frequency beta evaluated there, and the projection scale
// (2n+1) / (4 pi (g_n rho^n)^2) of every j.
struct Sphere {
    int p = 0, npt = 0, nj = 0;
    std::vector<double> pts, wts, psi, scale;
    Sphere(int p_, double beta, double rho, int ng) : p(p_) {
        nj = pw_nj(p);
        std::vector<double> x, w;
        gauss_legendre(ng, x, w);
        const int nf = 2 * ng;
        npt = ng * nf;
        pts.resize(size_t(3) * npt);
        wts.resize(npt);
        psi.resize(size_t(npt) * nj);
        for (int a = 0; a < ng; ++a)
            for (int b = 0; b < nf; ++b) {
                const int k = a * nf + b;
                const double ct = x[a], st = std::sqrt(std::max(0.0, 1 - ct * ct));
                const double ph = 2 * PI * b / nf;
                pts[3 * k] = rho * st * std::cos(ph);
                pts[3 * k + 1] = rho * st * std::sin(ph);
                pts[3 * k + 2] = rho * ct;
                wts[k] = w[a] * 2 * PI / nf;
                basis(p, beta, &pts[3 * k], &psi[size_t(k) * nj]);
            }
        double g[PW_MAXP + 2];
        pw_g<double, PW_MAXP>(beta * beta * rho * rho, g);
        scale.resize(nj);
        for (int n = 0; n <= p; ++n) {
            const double gn = g[n] * std::pow(rho, n);
            for (int j = n * n; j < (n + 1) * (n + 1); ++j) scale[j] = (2 * n + 1) / (4 * PI * gn * gn);
        }
    }
    // coefficients of f (values at the points, complex) in the basis
    template <typename V>
    void project(const std::vector<V>& f, std::vector<V>& c) const {
        c.assign(nj, V(0));
        for (int k = 0; k < npt; ++k) {
            const V fw = f[k] * wts[k];
            const double* ps = &psi[size_t(k) * nj];
            for (int j = 0; j < nj; ++j) c[j] += fw * ps[j];
        }
        for (int j = 0; j < nj; ++j) c[j] *= scale[j];
    }
};

int pad(int v, int q) { return (v + q - 1) / q * q; }

// reflection parity of basis function j under flipping axis a
double parity(int j, int a) {
    const int n = int(std::sqrt(double(j)) + 1e-9);
    const int r = j - n * n;
    const int m = (r + 1) / 2;
    const int part = r == 0 ? 0 : (r - 1) % 2;  // 0 cos, 1 sin
    if (a == 0) return ((m + part) % 2) ? -1.0 : 1.0;
    if (a == 1) return part ? -1.0 : 1.0;
    return ((n + m) % 2) ? -1.0 : 1.0;
}

int n_of(int j) { return int(std::sqrt(double(j)) + 1e-9); }
int m_of(int j) {
    const int n = n_of(j);
    return (j - n * n + 1) / 2;
}

// rotation matrices D^(d): psi_j(R_d^T s) = sum_j' D[j][j'] psi_j'(s) (per n)
std::vector<double> rotation_matrices(int p) {
    const int nj = pw_nj(p);
    double rot[6][3][3];
    pw_direction_frames(rot);
    Sphere sp(p, 0.0, 1.0, p + 4);
    std::vector<double> D(size_t(6) * nj * nj, 0.0);
    std::vector<double> vals(nj), f(sp.npt), c;
    for (int d = 0; d < 6; ++d) {
        // values psi_j(R^T s) at every point, projected for every j
        std::vector<double> pv(size_t(sp.npt) * nj);
        for (int k = 0; k < sp.npt; ++k) {
            const double* s = &sp.pts[3 * k];
            double l[3];
            for (int a = 0; a < 3; ++a) l[a] = rot[d][0][a] * s[0] + rot[d][1][a] * s[1] + rot[d][2][a] * s[2];
            basis(p, 0.0, l, &pv[size_t(k) * nj]);
        }
        for (int j = 0; j < nj; ++j) {
            for (int k = 0; k < sp.npt; ++k) f[k] = pv[size_t(k) * nj + j];
            sp.project(f, c);
            const int n = n_of(j);
            for (int jj = n * n; jj < (n + 1) * (n + 1); ++jj) D[(size_t(d) * nj + j) * nj + jj] = c[jj];
        }
    }
    return D;
}

}  // namespace

void pw_basis(int p, double beta, const double v[3], double* psi, double (*grad)[3]) {
    if (p < 0 || p > PW_MAXP) throw std::runtime_error("plane-wave order out of range");
    basis(p, beta, v, psi, grad);
}

void pw_direction_frames(double rot[6][3][3]) {
    // columns: local x, y, z axes (global components); rot[d][row][col]
    const double ax[6][3][3] = {
        {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},     // up:    x, y, z
        {{1, 0, 0}, {0, -1, 0}, {0, 0, -1}},   // down:  x, -y, -z
        {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}},     // north: z, x, y
        {{0, 0, 1}, {-1, 0, 0}, {0, -1, 0}},   // south: z, -x, -y
        {{0, 1, 0}, {0, 0, 1}, {1, 0, 0}},     // east:  y, z, x
        {{0, 1, 0}, {0, 0, -1}, {-1, 0, 0}},   // west:  y, -z, -x
    };
    for (int d = 0; d < 6; ++d)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) rot[d][r][c] = ax[d][c][r];
}

int pw_direction(int tx, int ty, int tz) {
    if (tz >= 2) return 0;
    if (tz <= -2) return 1;
    if (ty >= 2) return 2;
    if (ty <= -2) return 3;
    if (tx >= 2) return 4;
    if (tx <= -2) return 5;
    return -1;
}

PwQuad pw_quadrature(int nlam, double beta) {
    const double *T, *W;
    const int* NP;
    if (nlam == 9) {
        T = YR9_T; W = YR9_W; NP = YR9_NPHYS;
    } else if (nlam == 18) {
        T = YR18_T; W = YR18_W; NP = YR18_NPHYS;
    } else {
        throw std::runtime_error("plane-wave quadrature: 9 or 18 nodes");
    }
    PwQuad q;
    for (int k = 0; k < nlam; ++k) {
        const double w2 = std::sqrt(T[k] * (T[k] + 2 * beta));
        // angles: at least those of every Laplace node up to the first one
        // whose frequency reaches w2 (the screened wave oscillates faster)
        int M = NP[k];
        for (int j = k + 1; j < nlam; ++j) {
            if (T[j - 1] >= w2) break;
            M = std::max(M, NP[j]);
        }
        M += M % 2;
        q.u.push_back(T[k] + beta);
        q.w.push_back(w2);
        q.wt.push_back(W[k] / M);
        q.nhalf.push_back(M / 2);
        for (int a = 0; a < M / 2; ++a) {
            q.alpha.push_back(2 * PI * a / M);
            q.node.push_back(k);
        }
    }
    q.nw = int(q.alpha.size());
    return q;
}

namespace {

// tables of one level (see PwLevelTables); A7 for the M2M of this level's
// children is built by the caller
void level_waves(int p, int jp, int nwp, double beta, int nlam, const std::vector<double>& D,
                 PwLevelTables& t) {
    const int nj = pw_nj(p);
    t.q = pw_quadrature(nlam, beta);
    const PwQuad& q = t.q;
    const int nw = q.nw;
    const int nlam_ = int(q.u.size());
    // B_j(k, 0): coefficients of e^{u s_z - i w s_x}
    const int ng = 2 * p + 40;
    Sphere sp(p, beta, 0.75, ng);
    std::vector<std::vector<cd>> B0(nlam_);
    std::vector<cd> f(sp.npt);
    for (int k = 0; k < nlam_; ++k) {
        for (int i = 0; i < sp.npt; ++i) {
            const double* s = &sp.pts[3 * i];
            f[i] = std::exp(cd(q.u[k] * s[2], -q.w[k] * s[0]));
        }
        sp.project(f, B0[k]);
    }
    t.m2x.assign(size_t(jp) * nwp, 0.0);
    t.x2l.assign(size_t(nwp) * jp, 0.0);
    std::vector<cd> B(nj), E(nj);
    for (int wv = 0; wv < nw; ++wv) {
        const int k = q.node[wv];
        const double al = q.alpha[wv];
        // rotate about z by alpha
        for (int j = 0; j < nj; ++j) {
            const int n = n_of(j), m = m_of(j);
            if (m == 0) {
                B[j] = B0[k][j];
                continue;
            }
            const int jc = pw_idx(n, m, 0), js = pw_idx(n, m, 1);
            if (j != jc) continue;
            const double c = std::cos(m * al), s = std::sin(m * al);
            B[jc] = B0[k][jc] * c - B0[k][js] * s;
            B[js] = B0[k][jc] * s + B0[k][js] * c;
        }
        const double scale = q.wt[k];
        for (int d = 0; d < 6; ++d) {
            const double* Dd = &D[size_t(d) * nj * nj];
            const int cre = (d * nw + wv) * 2, cim = cre + 1;
            // M2X: W = scale sum_j B_j (D M)_j
            for (int jj = 0; jj < nj; ++jj) {
                cd acc = 0;
                const int n = n_of(jj);
                for (int j = n * n; j < (n + 1) * (n + 1); ++j) acc += B[j] * Dd[size_t(j) * nj + jj];
                t.m2x[size_t(jj) * nwp + cre] = scale * acc.real();
                t.m2x[size_t(jj) * nwp + cim] = scale * acc.imag();
                // X2L: E_j' = sum_j (-1)^n B_j D[j][j']; L += 2 Re(Wacc E)
                const double sg = (n % 2) ? -1.0 : 1.0;
                E[jj] = sg * acc;
            }
            for (int jj = 0; jj < nj; ++jj) {
                t.x2l[size_t(cre) * jp + jj] = 2 * E[jj].real();
                t.x2l[size_t(cim) * jp + jj] = -2 * E[jj].imag();
            }
        }
    }
    // X2X factors per local offset code
    t.shift.assign(size_t(98) * nw * 2, 0.0);
    for (int dz = 2; dz <= 3; ++dz)
        for (int dy = -3; dy <= 3; ++dy)
            for (int dx = -3; dx <= 3; ++dx) {
                const int code = (dz - 2) * 49 + (dy + 3) * 7 + (dx + 3);
                for (int wv = 0; wv < nw; ++wv) {
                    const int k = q.node[wv];
                    const double al = q.alpha[wv];
                    const cd e = std::exp(cd(-q.u[k] * dz, q.w[k] * (dx * std::cos(al) + dy * std::sin(al))));
                    t.shift[(size_t(code) * nw + wv) * 2] = e.real();
                    t.shift[(size_t(code) * nw + wv) * 2 + 1] = e.imag();
                }
            }
}

// A_7[j][j'] (child octant +,+,+): coefficients of psi^parent_j(d + s/2) in
// psi^child_j'(s), d = (1/4, 1/4, 1/4) in parent units
std::vector<double> shift_matrix(int p, double beta_child) {
    const int nj = pw_nj(p);
    const double beta_p = 2 * beta_child;
    // the integrand has degree <= 2p for Laplace and decays fast beyond it
    // for the screened kernel
    Sphere sp(p, beta_child, 1.0, p + 12);
    std::vector<double> A(size_t(nj) * nj);
    std::vector<double> vals(size_t(sp.npt) * nj), f(sp.npt), c;
    for (int i = 0; i < sp.npt; ++i) {
        const double v[3] = {0.25 + 0.5 * sp.pts[3 * i], 0.25 + 0.5 * sp.pts[3 * i + 1],
                             0.25 + 0.5 * sp.pts[3 * i + 2]};
        basis(p, beta_p, v, &vals[size_t(i) * nj]);
    }
    for (int j = 0; j < nj; ++j) {
        for (int i = 0; i < sp.npt; ++i) f[i] = vals[size_t(i) * nj + j];
        sp.project(f, c);
        for (int jj = 0; jj < nj; ++jj) A[size_t(j) * nj + jj] = c[jj];
    }
    return A;
}

void fill_shift_tables(int p, int jp, const std::vector<double>& A7, PwLevelTables& t) {
    const int nj = pw_nj(p);
    t.m2mT.assign(size_t(8) * jp * jp, 0.0);
    t.l2lT.assign(size_t(8) * jp * jp, 0.0);
    for (int o = 0; o < 8; ++o)
        for (int j = 0; j < nj; ++j)
            for (int jj = 0; jj < nj; ++jj) {
                double s = 1.0;
                for (int a = 0; a < 3; ++a)
                    if (!((o >> a) & 1)) s *= parity(j, a) * parity(jj, a);
                const double v = s * A7[size_t(j) * nj + jj];
                t.m2mT[(size_t(o) * jp + jj) * jp + j] = v;
                t.l2lT[(size_t(o) * jp + j) * jp + jj] = v;
            }
}

std::mutex g_rot_mu;
std::map<int, std::vector<double>> g_rot;

const std::vector<double>& rotations(int p) {
    std::lock_guard<std::mutex> lk(g_rot_mu);
    auto it = g_rot.find(p);
    if (it == g_rot.end()) it = g_rot.emplace(p, rotation_matrices(p)).first;
    return it->second;
}

// Process-wide caches of the level tables, keyed by everything they depend
// on (a plan of the same shape and frequency, or the Laplace kernel at any
// level, reuses them).  Bounded: cleared when large.
std::mutex g_cache_mu;
std::map<std::tuple<int, int, int, int, double>, PwLevelTables> g_waves;  // p, jp, nwp, nlam, beta
std::map<std::pair<int, double>, std::vector<double>> g_shift;            // p, beta_child

PwLevelTables cached_waves(int p, int jp, int nwp, int nlam, double beta) {
    const auto key = std::make_tuple(p, jp, nwp, nlam, beta);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_waves.find(key);
        if (it != g_waves.end()) return it->second;
    }
    PwLevelTables t;
    level_waves(p, jp, nwp, beta, nlam, rotations(p), t);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    if (g_waves.size() > 64) g_waves.clear();
    g_waves.emplace(key, t);
    return t;
}

std::vector<double> cached_shift(int p, double beta_child) {
    const auto key = std::make_pair(p, beta_child);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_shift.find(key);
        if (it != g_shift.end()) return it->second;
    }
    std::vector<double> A = shift_matrix(p, beta_child);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    if (g_shift.size() > 64) g_shift.clear();
    g_shift.emplace(key, A);
    return A;
}

}  // namespace

PwTableSet pw_build_tables(int p, int nlam, double kappa, double edge, int D) {
    if (p < 1 || p > PW_MAXP) throw std::runtime_error("plane-wave order out of range");
    PwTableSet ts;
    ts.p = p;
    ts.nlam = nlam;
    ts.jp = pad(pw_nj(p), 16);
    ts.lev.resize(D + 1);
    int nwmax = 0;
    for (int l = 0; l <= D; ++l) {
        const double h = edge / double(1L << l);
        ts.lev[l].h = h;
        ts.lev[l].beta = kappa * h;
        if (l >= 2) nwmax = std::max(nwmax, pw_quadrature(nlam, kappa * h).nw);
    }
    ts.nwp = pad(std::max(12 * nwmax, 64), 64);
    for (int l = 0; l <= D; ++l) {
        PwLevelTables& t = ts.lev[l];
        const double beta = t.beta, h = t.h;
        if (l >= 2) {
            t = cached_waves(p, ts.jp, ts.nwp, nlam, beta);
            t.beta = beta;
            t.h = h;
        }
        // M2M from this level's children (level l + 1) and L2L into them
        if (l < D) fill_shift_tables(p, ts.jp, cached_shift(p, kappa * h / 2), t);
    }
    return ts;
}

// ---------------------------------------------------------------------------
// self-test

void pw_selftest(int p, int nlam, double beta, double* err) {
    for (int i = 0; i < 5; ++i) err[i] = 0;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    auto kern = [&](double r) { return std::exp(-beta * r) / r; };
    // [0] quadrature of the kernel over the design region
    {
        const PwQuad q = pw_quadrature(nlam, beta);
        double emax = 0;
        for (int it = 0; it < 200; ++it) {
            const double z = 1 + 3 * (U(rng) + 0.5);
            const double x = 8 * U(rng), y = 8 * U(rng);
            const double r = std::sqrt(x * x + y * y + z * z);
            double s = 0;
            for (int wv = 0; wv < q.nw; ++wv) {
                const int k = q.node[wv];
                s += 2 * q.wt[k] * std::exp(-q.u[k] * z) *
                     std::cos(q.w[k] * (x * std::cos(q.alpha[wv]) + y * std::sin(q.alpha[wv])));
            }
            emax = std::max(emax, std::abs(s - kern(r)) / kern(1.0));
        }
        err[0] = emax;
    }
    const int nj = pw_nj(p);
    // three levels: root edge 4 so h = 4, 2, 1 at levels 0, 1, 2; beta is the
    // level-2 (box edge 1) frequency
    const PwTableSet ts = pw_build_tables(p, nlam, beta, 4.0, 3);
    const int jp = ts.jp, nwp = ts.nwp;
    // sources (charge + dipole) in a box of edge 1 at the origin
    const int NS = 20;
    std::vector<double> src(size_t(NS) * 7);
    for (int i = 0; i < NS; ++i) {
        for (int a = 0; a < 3; ++a) src[7 * i + a] = U(rng);
        src[7 * i + 3] = U(rng);
        for (int a = 0; a < 3; ++a) src[7 * i + 4 + a] = 0.3 * U(rng);
    }
    auto moments = [&](double b, const double* c, double hs, std::vector<double>& M) {
        // moments in box units of edge hs about c, frequency b (box units)
        M.assign(nj, 0.0);
        std::vector<double> ps(nj), gv(3 * size_t(nj));
        double(*g)[3] = reinterpret_cast<double(*)[3]>(gv.data());
        for (int i = 0; i < NS; ++i) {
            double v[3];
            for (int a = 0; a < 3; ++a) v[a] = (src[7 * i + a] - c[a]) / hs;
            basis(p, b, v, ps.data(), g);
            for (int j = 0; j < nj; ++j)
                M[j] += src[7 * i + 3] * ps[j] +
                        (src[7 * i + 4] * g[j][0] + src[7 * i + 5] * g[j][1] + src[7 * i + 6] * g[j][2]) / hs;
        }
    };
    auto direct = [&](const double* t, double& phi, double grad[3]) {
        phi = 0;
        grad[0] = grad[1] = grad[2] = 0;
        for (int i = 0; i < NS; ++i) {
            double dv[3];
            for (int a = 0; a < 3; ++a) dv[a] = t[a] - src[7 * i + a];
            const double r = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            const double G = kern(r);
            const double dG = -G * (beta + 1 / r) / r;  // dG/dr / r
            const double* dp = &src[7 * i + 4];
            const double q = src[7 * i + 3];
            const double dd = dp[0] * dv[0] + dp[1] * dv[1] + dp[2] * dv[2];
            // phi = q G(t - s) + d . grad_s G = q G - dG dd
            phi += q * G - dG * dd;
            // grad_t of (q G) = q dG dv ; of (-dG(r) dd): -(d2 dv (dd)/r + dG dp)
            const double d2G = G * ((beta * beta + 3 * beta / r + 3 / (r * r)) / (r * r));  // (d/dr (dG/dr / r)) / r
            for (int a = 0; a < 3; ++a) grad[a] += q * dG * dv[a] - (d2G * dd * dv[a] + dG * dp[a]);
        }
    };
    auto eval_local = [&](const std::vector<double>& L, double b, const double* tc, double ht, const double* t,
                          double& phi, double grad[3]) {
        std::vector<double> ps(nj), gv(3 * size_t(nj));
        double(*g)[3] = reinterpret_cast<double(*)[3]>(gv.data());
        double v[3];
        for (int a = 0; a < 3; ++a) v[a] = (t[a] - tc[a]) / ht;
        basis(p, b, v, ps.data(), g);
        phi = 0;
        grad[0] = grad[1] = grad[2] = 0;
        for (int j = 0; j < nj; ++j) {
            phi += L[j] * ps[j];
            for (int a = 0; a < 3; ++a) grad[a] += L[j] * g[j][a] / ht;
        }
    };
    // apply M2X (level lt), one direction's X2X and X2L
    auto m2x_x2l = [&](const PwLevelTables& t, const std::vector<double>& M, int d, int code) {
        const int nw = t.q.nw;
        std::vector<double> L(nj, 0.0);
        for (int wv = 0; wv < nw; ++wv) {
            const int cre = (d * nw + wv) * 2;
            double wr = 0, wi = 0;
            for (int j = 0; j < nj; ++j) {
                wr += M[j] * t.m2x[size_t(j) * nwp + cre];
                wi += M[j] * t.m2x[size_t(j) * nwp + cre + 1];
            }
            const double sr = t.shift[(size_t(code) * nw + wv) * 2], si = t.shift[(size_t(code) * nw + wv) * 2 + 1];
            const double ar = wr * sr - wi * si, ai = wr * si + wi * sr;
            for (int j = 0; j < nj; ++j) L[j] += ar * t.x2l[size_t(cre) * jp + j] + ai * t.x2l[size_t(cre + 1) * jp + j];
        }
        return L;
    };
    double rot[6][3][3];
    pw_direction_frames(rot);
    const int offs[6][3] = {{1, -2, 3}, {0, 3, -2}, {-3, 2, 1}, {2, -3, 0}, {3, 1, -1}, {-2, 0, 1}};
    // errors relative to the largest value of each quantity over the case
    double e1 = 0, e3 = 0, e2 = 0;
    for (int c = 0; c < 6; ++c) {
        const int* o = offs[c];
        const int d = pw_direction(o[0], o[1], o[2]);
        double tl[3];
        for (int a = 0; a < 3; ++a) tl[a] = rot[d][0][a] * o[0] + rot[d][1][a] * o[1] + rot[d][2][a] * o[2];
        const int code = (int(std::lround(tl[2])) - 2) * 49 + (int(std::lround(tl[1])) + 3) * 7 + int(std::lround(tl[0])) + 3;
        const PwLevelTables& t2 = ts.lev[2];
        std::vector<double> M;
        const double c0[3] = {0, 0, 0};
        moments(beta, c0, 1.0, M);
        const std::vector<double> L = m2x_x2l(t2, M, d, code);
        const double tc[3] = {double(o[0]), double(o[1]), double(o[2])};
        double pmax = 0, gmax = 0, pe = 0, ge = 0;
        for (int it = 0; it < 20; ++it) {
            double tp[3];
            for (int a = 0; a < 3; ++a) tp[a] = tc[a] + U(rng);
            double pd, gd[3], pf, gf[3];
            direct(tp, pd, gd);
            eval_local(L, beta, tc, 1.0, tp, pf, gf);
            pmax = std::max(pmax, std::abs(pd));
            pe = std::max(pe, std::abs(pf - pd));
            for (int a = 0; a < 3; ++a) {
                gmax = std::max(gmax, std::abs(gd[a]));
                ge = std::max(ge, std::abs(gf[a] - gd[a]));
            }
        }
        e1 = std::max(e1, pe / pmax);
        e3 = std::max(e3, ge / gmax);
        // through M2M and L2L: sources in child octant 5 (+x, -y, +z) of the
        // level-2 box (edge 1) at the origin, target in child octant 2 of the
        // level-2 box at offset o: child moments -> M2M -> waves at level 2
        // -> L2L -> child local expansion (level 3, edge 1/2)
        {
            const int oc = 5;
            const double ch[3] = {(oc & 1) ? 0.25 : -0.25, (oc & 2) ? 0.25 : -0.25, (oc & 4) ? 0.25 : -0.25};
            // move the sources into the child box: s -> ch + s/2
            std::vector<double> saved = src;
            for (int i = 0; i < NS; ++i)
                for (int a = 0; a < 3; ++a) src[7 * i + a] = ch[a] + 0.5 * saved[7 * i + a];
            std::vector<double> Mc;
            moments(beta / 2, ch, 0.5, Mc);
            // M2M with level 2's tables (children at level 3)
            const PwLevelTables& tp2 = ts.lev[2];
            std::vector<double> Mp(nj, 0.0);
            for (int j = 0; j < nj; ++j)
                for (int jj = 0; jj < nj; ++jj) Mp[j] += tp2.m2mT[(size_t(oc) * jp + jj) * jp + j] * Mc[jj];
            const std::vector<double> Lp = m2x_x2l(tp2, Mp, d, code);
            // L2L into the target parent's child octant 2
            const int ot = 2;
            const double tch[3] = {o[0] + ((ot & 1) ? 0.25 : -0.25), o[1] + ((ot & 2) ? 0.25 : -0.25),
                                   o[2] + ((ot & 4) ? 0.25 : -0.25)};
            std::vector<double> Lc(nj, 0.0);
            for (int j = 0; j < nj; ++j)
                for (int jj = 0; jj < nj; ++jj) Lc[jj] += Lp[j] * tp2.l2lT[(size_t(ot) * jp + j) * jp + jj];
            double pmax2 = 0, pe2 = 0;
            for (int it = 0; it < 20; ++it) {
                double tp[3];
                for (int a = 0; a < 3; ++a) tp[a] = tch[a] + 0.5 * U(rng);
                double pd, gd[3], pf, gf[3];
                direct(tp, pd, gd);
                eval_local(Lc, beta / 2, tch, 0.5, tp, pf, gf);
                pmax2 = std::max(pmax2, std::abs(pd));
                pe2 = std::max(pe2, std::abs(pf - pd));
            }
            e2 = std::max(e2, pe2 / pmax2);
            src = saved;
        }
    }
    err[1] = e1;
    err[2] = e2;
    err[3] = e3;
    // [4] rotation orthogonality
    const std::vector<double>& Dr = rotations(p);
    double e4 = 0;
    for (int d = 0; d < 6; ++d)
        for (int j = 0; j < nj; ++j)
            for (int jj = 0; jj < nj; ++jj) {
                double s = 0;
                for (int k = 0; k < nj; ++k) s += Dr[(size_t(d) * nj + j) * nj + k] * Dr[(size_t(d) * nj + jj) * nj + k];
                e4 = std::max(e4, std::abs(s - (j == jj ? 1.0 : 0.0)));
            }
    err[4] = e4;
}

}  // namespace beepdev
