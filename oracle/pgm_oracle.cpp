/* pgm_oracle: test infrastructure only (the checker of the HIP engine; never on the product path).
 *
 * FP64 CPU restatement of the reference engine core/pgm_core.cpp and its numerics core/pgm_math.h,
 * written from the model in docs/spec.md.  Every function cites the reference lines it follows.  What
 * differs from the reference, by design:
 *   - the direct-space pairs come from a cell list instead of the O(N^2) loop (core/pgm_core.cpp:64-87);
 *     the pair set and the pair order (i < j, sorted by i then j) are the same, so sums run in the same
 *     order;
 *   - the 3D FFT is an in-house mixed-radix FFT instead of pocketfft (core/pgm_core.cpp:213,237).
 */
#include "pgm_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <deque>
#include <memory>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

typedef std::array<double, 3> V3;
typedef std::complex<double> cd;

constexpr double kSqrtPi = 1.7724538509055160273;
constexpr double kPi = 3.14159265358979323846;
constexpr double kMinPolarizability = 1e-9;   // core/pgm_constants.h:9

thread_local std::string g_err;

inline double dot3(const V3& a, const V3& b) { return a[0]*b[0]+a[1]*b[1]+a[2]*b[2]; }

// ------------------------------------------------------------------------------------------------ numerics

/* Radial kernels B_n of erf(a r)/r, B_{n+1} = -(1/r) dB_n/dr (core/pgm_math.h:31-57): the 15-term series in
 * s = (a r)^2 below s = 1, the upward recursion above. */
void erfKernels(double a, double r, int nmax, double* B) {
    const double x = a*r, s = x*x, pre = 2.0*a/kSqrtPi, twoA2 = 2.0*a*a;
    if (s < 1.0) {
        double term[15];
        term[0] = 1.0;
        for (int m = 1; m < 15; m++)
            term[m] = -term[m-1]*s/m;
        double scale = pre;
        for (int k = 0; k < nmax; k++, scale *= twoA2) {
            double acc = 0.0;
            for (int m = 14; m >= 0; m--)
                acc += term[m]/(2*m+2*k+1);
            B[k] = scale*acc;
        }
        return;
    }
    const double g = pre*std::exp(-s), inv2 = 1.0/(r*r);
    B[0] = std::erf(x)/r;
    double scale = 1.0;
    for (int k = 1; k < nmax; k++, scale *= twoA2)
        B[k] = ((2*k-1)*B[k-1]-scale*g)*inv2;
}

/* B_n of erfc(a r)/r: all terms positive (core/pgm_math.h:63-71). */
void erfcKernels(double a, double r, int nmax, double* B) {
    const double x = a*r, g = (2.0/kSqrtPi)*a*std::exp(-x*x), inv2 = 1.0/(r*r), twoA2 = 2.0*a*a;
    B[0] = std::erfc(x)/r;
    double scale = 1.0;
    for (int k = 1; k < nmax; k++, scale *= twoA2)
        B[k] = ((2*k-1)*B[k-1]+scale*g)*inv2;
}

/* G_n = B_n[erf(a r)/r] - B_n[erf(beta r)/r]: the series difference for a r < 1, else the erfc form
 * Bc(beta) - Bc(a) (core/pgm_math.h:78-92). */
void pairKernels(double a, double beta, double r, int nmax, double* G) {
    double other[4];
    if (a*r < 1.0) {
        erfKernels(a, r, nmax, G);
        erfKernels(beta, r, nmax, other);
    }
    else {
        erfcKernels(beta, r, nmax, G);
        erfcKernels(a, r, nmax, other);
    }
    for (int k = 0; k < nmax; k++)
        G[k] -= other[k];
}

/* |sum_k M_p(k+1) e^{2 pi i m k/K}|^2 with zeros replaced by the neighbour mean (core/pgm_math.h:98-126). */
std::vector<double> bsplineModuli(int K, int p) {
    // cardinal B-spline M_p at the integers 1..p-1 by the Cox-de Boor recursion at fraction 0
    std::vector<double> prev(p+1, 0.0), cur(p+1, 0.0);
    prev[0] = 1.0;   // M_1(0 + j), j = 0
    for (int k = 2; k <= p; k++) {
        for (int j = 0; j < k; j++) {
            const double hi = (j < k-1 ? prev[j] : 0.0), lo = (j > 0 ? prev[j-1] : 0.0);
            cur[j] = (j*hi+(k-j)*lo)/(k-1);
        }
        for (int j = 0; j < k; j++)
            prev[j] = cur[j];
    }
    std::vector<double> coef(K, 0.0);
    for (int k = 0; k+1 < p && k < K; k++)
        coef[k] = prev[k+1];
    std::vector<double> mod(K);
    for (int m = 0; m < K; m++) {
        double re = 0, im = 0;
        for (int k = 0; k < K; k++) {
            const double ang = 2*kPi*m*k/K;
            re += coef[k]*std::cos(ang);
            im += coef[k]*std::sin(ang);
        }
        mod[m] = re*re+im*im;
    }
    for (int m = 0; m < K; m++)
        if (mod[m] < 1e-7)
            mod[m] = 0.5*(mod[(m+K-1)%K]+mod[(m+1)%K]);
    return mod;
}

/* beta with erfc(beta rc) = rtol by doubling then bisection (core/pgm_math.h:129-146). */
double ewaldCoefficient(double rc, double rtol) {
    double beta = 5.0, lo = 0.0;
    int steps = 0;
    do {
        steps++;
        beta *= 2;
    } while (std::erfc(beta*rc) > rtol);
    double hi = beta;
    for (int i = 0; i < steps+60; i++) {
        beta = 0.5*(lo+hi);
        if (std::erfc(beta*rc) > rtol)
            lo = beta;
        else
            hi = beta;
    }
    return beta;
}

/* Reduced-triclinic minimum image: c, then b, then a (core/pgm_math.h:153-160). */
inline void minImage(V3& v, const double box[3][3]) {
    for (int c = 2; c >= 0; c--) {
        const double k = std::round(v[c]/box[c][c]);
        v[0] -= k*box[c][0];
        v[1] -= k*box[c][1];
        v[2] -= k*box[c][2];
    }
}

// ------------------------------------------------------------------------------------------------ GVDW

/* G(y) = F(y)/y^6, F = A^2 + g^2/2, A = erf y - (2y/sqrt(pi))(1 + 2y^2/3) e^{-y^2}, g = (4y^3/(3 sqrt(pi)))
 * e^{-y^2}, and H(y) = G'(y)/y (core/gvdw_kernel.h:30-55): below y = 0.6 the Taylor series of G in y^2 (its
 * coefficients are those of (9/2) alpha(s)^2/s^2 + e^{-2s} with A = (2/sqrt(pi)) y alpha(y^2); 15 terms),
 * the closed form above. */
void gvdwG(double y, double& G, double& H) {
    static const double c[15] = {1.0, -2.0, 58.0/25, -188.0/105, 2222.0/2205, -1532.0/3465, 64244.0/405405,
                                 -162056.0/3378375, 1587734.0/126351225, -211108.0/72747675,
                                 11902564.0/19860115275.0, -611992.0/5465775315.0,
                                 755028172.0/39529267903125.0, -213335384.0/71152682225625.0,
                                 15291334888.0/35078272337233125.0};
    const double c0 = 8.0/(9.0*kPi), s = y*y;
    if (y < 0.6) {
        double g = 0, h = 0;
        for (int k = 14; k >= 0; k--) {
            g = g*s+c[k];
            if (k > 0)
                h = h*s+2.0*k*c[k];
        }
        G = c0*g;
        H = c0*h;
        return;
    }
    const double e = std::exp(-s), isp = 1.0/kSqrtPi;
    const double A = std::erf(y)-2.0*isp*y*(1.0+2.0*s/3.0)*e, g = (4.0/3.0)*isp*s*y*e;
    const double dA = (8.0/3.0)*isp*s*s*e, dg = 4.0*isp*s*(1.0-2.0*s/3.0)*e;
    const double F = A*A+0.5*g*g, dF = 2.0*A*dA+g*dg, y6 = s*s*s;
    G = F/y6;
    H = (dF/y6-6.0*F/(y6*y))/y;
}

/* U(r) = A exp(-B y^2) (Gaussian) or A exp(-B y) (Slater) - C6 b^6 G(y), y = b r; dEdR = -(1/r) dU/dr
 * (core/gvdw_kernel.h:58-77); the Slater force at r = 0 is 0. */
double gvdwPair(double r, double bij, double A, double C6, double B, int rep, double& dEdR) {
    const double y = bij*r, b2 = bij*bij, b6 = b2*b2*b2;
    double erep, drep;
    if (rep == 0) {
        erep = A*std::exp(-B*y*y);
        drep = -2.0*B*b2*erep;
    }
    else {
        erep = A*std::exp(-B*y);
        drep = r > 0 ? -B*bij*erep/r : 0.0;
    }
    double G, H;
    gvdwG(y, G, H);
    dEdR = C6*b6*b2*H-drep;
    return erep-C6*b6*G;
}

// ------------------------------------------------------------------------------------------------ FFT

/* Mixed-radix complex FFT (Stockham autosort over the prime factors of n, radix 4 first), unnormalized,
 * sign -1 forward, +1 backward.  Plans are cached per (n, sign). */
class Fft1d {
public:
    Fft1d(int n, int sign) : n_(n), work_(n) {
        int m = n;
        while (m%4 == 0) {
            radix_.push_back(4);
            m /= 4;
        }
        for (int p = 2; p <= m; p++)
            while (m%p == 0) {
                radix_.push_back(p);
                m /= p;
            }
        int ns = 1;
        for (int R : radix_) {
            stageTw_.emplace_back(R*ns);
            for (int r = 0; r < R; r++)
                for (int k = 0; k < ns; k++) {
                    const double ang = sign*2.0*kPi*r*k/(double) (ns*R);
                    stageTw_.back()[r*ns+k] = cd(std::cos(ang), std::sin(ang));
                }
            roots_.emplace_back(R);
            for (int r = 0; r < R; r++) {
                const double ang = sign*2.0*kPi*r/(double) R;
                roots_.back()[r] = cd(std::cos(ang), std::sin(ang));
            }
            ns *= R;
        }
    }
    void run(cd* data) {
        cd* src = data;
        cd* dst = work_.data();
        int ns = 1;
        cd v[64];
        for (size_t st = 0; st < radix_.size(); st++) {
            const int R = radix_[st], m = n_/R;
            const std::vector<cd>& tw = stageTw_[st];
            const std::vector<cd>& root = roots_[st];
            for (int j = 0; j < m; j++) {
                const int k = j%ns;
                for (int r = 0; r < R; r++)
                    v[r] = src[j+r*m]*tw[r*ns+k];
                cd* out = dst+(j/ns)*ns*R+k;
                if (R == 2) {
                    out[0] = v[0]+v[1];
                    out[ns] = v[0]-v[1];
                }
                else if (R == 4) {
                    const cd a = v[0]+v[2], b = v[0]-v[2], c = v[1]+v[3], d = (v[1]-v[3])*root[1];
                    out[0] = a+c;
                    out[ns] = b+d;
                    out[2*ns] = a-c;
                    out[3*ns] = b-d;
                }
                else
                    for (int q = 0; q < R; q++) {
                        cd acc = 0;
                        for (int r = 0; r < R; r++)
                            acc += v[r]*root[(r*q)%R];
                        out[q*ns] = acc;
                    }
            }
            std::swap(src, dst);
            ns *= R;
        }
        if (src != data)
            std::copy(src, src+n_, data);
    }

private:
    int n_;
    std::vector<cd> work_;
    std::vector<int> radix_;
    std::vector<std::vector<cd> > stageTw_, roots_;
};

Fft1d& fftPlan(int n, int sign) {
    thread_local std::vector<std::pair<std::pair<int, int>, std::unique_ptr<Fft1d> > > cache;
    for (auto& e : cache)
        if (e.first.first == n && e.first.second == sign)
            return *e.second;
    if (n > 0) {
        int m = n;
        for (int p = 2; p <= m; p++)
            while (m%p == 0) {
                if (p > 64)
                    throw std::invalid_argument("pgm_oracle: FFT length has a prime factor above 64");
                m /= p;
            }
    }
    cache.emplace_back(std::make_pair(n, sign), std::unique_ptr<Fft1d>(new Fft1d(n, sign)));
    return *cache.back().second;
}

/* In-place 3D complex FFT of a row-major K1 x K2 x K3 array along all three axes. */
void fft3d(std::vector<cd>& a, int K1, int K2, int K3, int sign) {
    const int K[3] = {K1, K2, K3};
    const size_t strides[3] = {(size_t) K2*K3, (size_t) K3, 1};
    for (int ax = 2; ax >= 0; ax--) {
        Fft1d& f = fftPlan(K[ax], sign);
        std::vector<cd> line(K[ax]);
        const int o1 = (ax == 0 ? 1 : 0), o2 = (ax == 2 ? 1 : 2);
        for (int i = 0; i < K[o1]; i++)
            for (int j = 0; j < K[o2]; j++) {
                const size_t base = i*strides[o1]+j*strides[o2];
                if (ax == 2) {
                    f.run(&a[base]);
                    continue;
                }
                for (int k = 0; k < K[ax]; k++)
                    line[k] = a[base+k*strides[ax]];
                f.run(line.data());
                for (int k = 0; k < K[ax]; k++)
                    a[base+k*strides[ax]] = line[k];
            }
    }
}

/* Forward real-to-complex 3D FFT: half spectrum K1 x K2 x (K3/2+1), as pocketfft::r2c with FORWARD. */
void r2c3d(const double* in, cd* out, int K1, int K2, int K3) {
    const int K3h = K3/2+1;
    std::vector<cd> a((size_t) K1*K2*K3);
    for (size_t i = 0; i < a.size(); i++)
        a[i] = cd(in[i], 0.0);
    fft3d(a, K1, K2, K3, -1);
    for (int i = 0; i < K1; i++)
        for (int j = 0; j < K2; j++)
            for (int k = 0; k < K3h; k++)
                out[((size_t) i*K2+j)*K3h+k] = a[((size_t) i*K2+j)*K3+k];
}

/* Backward complex-to-real 3D FFT of a Hermitian half spectrum (the other half by conjugate symmetry). */
void c2r3d(const cd* in, double* out, int K1, int K2, int K3) {
    const int K3h = K3/2+1;
    std::vector<cd> a((size_t) K1*K2*K3);
    for (int i = 0; i < K1; i++)
        for (int j = 0; j < K2; j++)
            for (int k = 0; k < K3; k++) {
                cd v;
                if (k < K3h)
                    v = in[((size_t) i*K2+j)*K3h+k];
                else
                    v = std::conj(in[((size_t) ((K1-i)%K1)*K2+(K2-j)%K2)*K3h+(K3-k)]);
                a[((size_t) i*K2+j)*K3+k] = v;
            }
    fft3d(a, K1, K2, K3, +1);
    for (size_t i = 0; i < a.size(); i++)
        out[i] = a[i].real();
}

// ------------------------------------------------------------------------------------------------ engine

struct Engine {
    orc_settings s;
    int n = 0;
    std::vector<double> q, radius, alpha;
    std::vector<int> covAtom, covPartner;
    std::vector<double> covStrength;
    std::vector<V3> perm, mu;
    std::deque<std::vector<V3> > history;
    std::vector<V3> lastPos;
    bool havePos = false;
    long long historyKey = -1, lastKey = -1;
    int lastIterations = 0;
    double lastResidual = 0;
    double terms[5] = {0, 0, 0, 0, 0};
    double strain[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // geometry
    std::vector<V3> pos;
    double box[3][3] = {{0}}, recip[3][3] = {{0}};   // recip[k][a] = (H^-1)_{k a}
    double volume = 0;
    struct Pair {
        int i, j;
        V3 x;
        double G[4];
    };
    std::vector<Pair> pairs;
    // CG preconditioner: the short-range one, M = D + D T_s D over the pairs within preRadius (0: Jacobi, M = D),
    // at most kPreCap such pairs per atom or Jacobi everywhere (core/cuda/pgm_cuda.cu:162,1526-1588); Jacobi with
    // the self term of the diagonal, alpha/(1 - alpha c_self) (:1461-1468,1758-1760).  Both are options of the
    // reference's CUDA engine only; core/pgm_core.cpp is Jacobi.
    static constexpr int kPreCap = 48;
    double preRadius = 0;
    bool jacobiSelf = false;
    bool preOn = false;
    std::vector<std::vector<int> > preList;   // per atom: its short-range pairs (indices into pairs)
    // PME
    std::vector<int> gidx[3];
    std::vector<double> th[3], dth[3], d2th[3];
    std::vector<double> influence;

    void setTopology(const double* q_, const double* r_, const double* a_, int ncov, const int* ca, const int* cp,
                     const double* cs) {
        q.assign(q_, q_+n);
        radius.assign(r_, r_+n);
        alpha.assign(a_, a_+n);
        for (double& v : alpha)   // core/pgm_core.cpp:17-21
            if (v <= kMinPolarizability)
                v = 0.0;
        covAtom.assign(ca, ca+ncov);
        covPartner.assign(cp, cp+ncov);
        covStrength.assign(cs, cs+ncov);
        for (int m = 0; m < ncov; m++)
            if (covAtom[m] < 0 || covAtom[m] >= n || covPartner[m] < 0 || covPartner[m] >= n)
                throw std::invalid_argument("pgm_oracle: covalent dipole atom index out of range");
    }

    // p_i += c unit(r_j - r_i), minimum image (core/pgm_core.cpp:51-62)
    void permDipoles() {
        perm.assign(n, V3{0, 0, 0});
        for (size_t m = 0; m < covAtom.size(); m++) {
            const int i = covAtom[m], j = covPartner[m];
            V3 v = {pos[j][0]-pos[i][0], pos[j][1]-pos[i][1], pos[j][2]-pos[i][2]};
            if (s.periodic)
                minImage(v, box);
            const double len = std::sqrt(dot3(v, v));
            for (int k = 0; k < 3; k++)
                perm[i][k] += covStrength[m]*v[k]/len;
        }
    }

    void addPair(int i, int j, const V3& x) {
        Pair p;
        p.i = i;
        p.j = j;
        p.x = x;
        const double r = std::sqrt(dot3(x, x));
        const double a = 1.0/std::sqrt(2.0*(radius[i]*radius[i]+radius[j]*radius[j]));
        if (s.periodic)
            pairKernels(a, s.beta, r, 4, p.G);
        else
            erfKernels(a, r, 4, p.G);
        pairs.push_back(p);
    }

    /* The pairs i < j within the cutoff (all pairs without a box), in the order of the reference's double
     * loop (core/pgm_core.cpp:64-87).  Periodic: candidates from fractional cells at least rc/2 wide with a
     * +-2 stencil; the test itself is the reference's (minimum image of r_i - r_j, r^2 < rc^2). */
    void buildPairs() {
        pairs.clear();
        if (!s.periodic) {
            for (int i = 0; i < n; i++)
                for (int j = i+1; j < n; j++)
                    addPair(i, j, V3{pos[i][0]-pos[j][0], pos[i][1]-pos[j][1], pos[i][2]-pos[j][2]});
            return;
        }
        const double rc2 = s.cutoff*s.cutoff;
        int nc[3];
        for (int a = 0; a < 3; a++) {
            const double w = 1.0/std::sqrt(recip[0][a]*recip[0][a]+recip[1][a]*recip[1][a]+recip[2][a]*recip[2][a]);
            nc[a] = std::max(1, (int) std::floor(w/(0.5*s.cutoff)));
        }
        std::vector<int> cellOf(n);
        const size_t ncell = (size_t) nc[0]*nc[1]*nc[2];
        std::vector<std::vector<int> > members(ncell);
        std::vector<std::array<int, 3> > cidx(n);
        for (int i = 0; i < n; i++) {
            int c[3];
            for (int a = 0; a < 3; a++) {
                double f = pos[i][0]*recip[0][a]+pos[i][1]*recip[1][a]+pos[i][2]*recip[2][a];
                f -= std::floor(f);
                c[a] = std::min(nc[a]-1, std::max(0, (int) std::floor(f*nc[a])));
                cidx[i][a] = c[a];
            }
            const size_t id = ((size_t) c[0]*nc[1]+c[1])*nc[2]+c[2];
            members[id].push_back(i);
        }
        std::vector<int> offs[3];
        for (int a = 0; a < 3; a++) {
            if (nc[a] >= 5)
                offs[a] = {-2, -1, 0, 1, 2};
            else
                for (int o = 0; o < nc[a]; o++)
                    offs[a].push_back(o);
        }
        std::vector<int> cand;
        for (int i = 0; i < n; i++) {
            cand.clear();
            for (int o0 : offs[0])
                for (int o1 : offs[1])
                    for (int o2 : offs[2]) {
                        const int c0 = ((cidx[i][0]+o0)%nc[0]+nc[0])%nc[0];
                        const int c1 = ((cidx[i][1]+o1)%nc[1]+nc[1])%nc[1];
                        const int c2 = ((cidx[i][2]+o2)%nc[2]+nc[2])%nc[2];
                        for (int j : members[((size_t) c0*nc[1]+c1)*nc[2]+c2])
                            if (j > i)
                                cand.push_back(j);
                    }
            std::sort(cand.begin(), cand.end());
            for (int j : cand) {
                V3 x = {pos[i][0]-pos[j][0], pos[i][1]-pos[j][1], pos[i][2]-pos[j][2]};
                minImage(x, box);
                if (dot3(x, x) >= rc2)
                    continue;
                addPair(i, j, x);
            }
        }
    }

    // B-spline weights on grid points (base - j) mod K (core/pgm_core.cpp:123-153)
    void setupPme() {
        const int p = s.order;
        for (int a = 0; a < 3; a++) {
            gidx[a].assign(n, 0);
            th[a].assign((size_t) n*p, 0.0);
            dth[a].assign((size_t) n*p, 0.0);
            d2th[a].assign((size_t) n*p, 0.0);
        }
        // spl[k][j] = M_k(fr + j), j = 0..k-1
        std::vector<std::vector<double> > spl(p+1, std::vector<double>(p+2, 0.0));
        auto at = [&](int k, int j) { return (j >= 0 && j < k) ? spl[k][j] : 0.0; };
        for (int i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                double t = pos[i][0]*recip[0][a]+pos[i][1]*recip[1][a]+pos[i][2]*recip[2][a];
                t -= std::floor(t);
                const double w = s.grid[a]*t;
                const int base = (int) std::floor(w);
                const double fr = w-base;
                spl[1][0] = 1.0;
                for (int k = 2; k <= p; k++)
                    for (int j = 0; j < k; j++)
                        spl[k][j] = ((fr+j)*at(k-1, j)+(k-fr-j)*at(k-1, j-1))/(k-1);
                for (int j = 0; j < p; j++) {
                    th[a][(size_t) i*p+j] = spl[p][j];
                    dth[a][(size_t) i*p+j] = at(p-1, j)-at(p-1, j-1);
                    d2th[a][(size_t) i*p+j] = at(p-2, j)-2*at(p-2, j-1)+at(p-2, j-2);
                }
                gidx[a][i] = base;
            }
    }

    // G(m) = exp(-pi^2 m^2/beta^2)/(pi V m^2)/(b1 b2 b3) on the half spectrum (core/pgm_core.cpp:155-177)
    void computeInfluence() {
        const int K1 = s.grid[0], K2 = s.grid[1], K3 = s.grid[2], K3h = K3/2+1;
        const std::vector<double> b1 = bsplineModuli(K1, s.order), b2 = bsplineModuli(K2, s.order),
                                  b3 = bsplineModuli(K3, s.order);
        influence.assign((size_t) K1*K2*K3h, 0.0);
        const double fac = kPi*kPi/(s.beta*s.beta);
        for (int m1 = 0; m1 < K1; m1++) {
            const int f1 = (m1 <= K1/2 ? m1 : m1-K1);
            for (int m2 = 0; m2 < K2; m2++) {
                const int f2 = (m2 <= K2/2 ? m2 : m2-K2);
                for (int m3 = 0; m3 < K3h; m3++) {
                    if (f1 == 0 && f2 == 0 && m3 == 0)
                        continue;
                    V3 mv;
                    for (int k = 0; k < 3; k++)
                        mv[k] = f1*recip[k][0]+f2*recip[k][1]+m3*recip[k][2];
                    const double m2s = dot3(mv, mv);
                    influence[((size_t) m1*K2+m2)*K3h+m3] =
                        std::exp(-fac*m2s)/(kPi*volume*m2s)/(b1[m1]*b2[m2]*b3[m3]);
                }
            }
        }
    }

    void setGeometry(const double* p, const double* box9) {
        pos.resize(n);
        for (int i = 0; i < n; i++)
            pos[i] = V3{p[3*i], p[3*i+1], p[3*i+2]};
        if (s.periodic) {
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    box[r][c] = box9[3*r+c];
            // inverse of H (rows = box vectors) by cofactors (core/pgm_core.cpp:95-111)
            const double(*H)[3] = box;
            const double det = H[0][0]*(H[1][1]*H[2][2]-H[1][2]*H[2][1])-H[0][1]*(H[1][0]*H[2][2]-H[1][2]*H[2][0])
                               +H[0][2]*(H[1][0]*H[2][1]-H[1][1]*H[2][0]);
            if (det == 0)
                throw std::invalid_argument("pgm_oracle: singular box");
            volume = std::fabs(det);
            recip[0][0] = (H[1][1]*H[2][2]-H[1][2]*H[2][1])/det;
            recip[0][1] = (H[0][2]*H[2][1]-H[0][1]*H[2][2])/det;
            recip[0][2] = (H[0][1]*H[1][2]-H[0][2]*H[1][1])/det;
            recip[1][0] = (H[1][2]*H[2][0]-H[1][0]*H[2][2])/det;
            recip[1][1] = (H[0][0]*H[2][2]-H[0][2]*H[2][0])/det;
            recip[1][2] = (H[0][2]*H[1][0]-H[0][0]*H[1][2])/det;
            recip[2][0] = (H[1][0]*H[2][1]-H[1][1]*H[2][0])/det;
            recip[2][1] = (H[0][1]*H[2][0]-H[0][0]*H[2][1])/det;
            recip[2][2] = (H[0][0]*H[1][1]-H[0][1]*H[1][0])/det;
        }
        permDipoles();
        buildPairs();
        if (s.periodic) {
            setupPme();
            computeInfluence();
        }
    }

    // e_a = K_a d . recip[:, a] (core/pgm_core.cpp:179-185)
    std::vector<V3> fracDipoles(const std::vector<V3>& d) const {
        std::vector<V3> e(n);
        for (int i = 0; i < n; i++)
            for (int a = 0; a < 3; a++)
                e[i][a] = s.grid[a]*(d[i][0]*recip[0][a]+d[i][1]*recip[1][a]+d[i][2]*recip[2][a]);
        return e;
    }

    inline size_t gridPoint(int i, int j1, int j2, int j3) const {
        const int K1 = s.grid[0], K2 = s.grid[1], K3 = s.grid[2];
        const int g1 = ((gidx[0][i]-j1)%K1+K1)%K1, g2 = ((gidx[1][i]-j2)%K2+K2)%K2, g3 = ((gidx[2][i]-j3)%K3+K3)%K3;
        return ((size_t) g1*K2+g2)*K3+g3;
    }

    /* Spread charges and dipoles, FFT, energy (and k-space strain), convolve, inverse FFT: the potential
     * phi on the grid (core/pgm_core.cpp:187-239). */
    double pmePotential(const std::vector<double>* qv, const std::vector<V3>& d, std::vector<double>& phi,
                        double* S) {
        const int K1 = s.grid[0], K2 = s.grid[1], K3 = s.grid[2], K3h = K3/2+1, p = s.order;
        std::vector<double> Q((size_t) K1*K2*K3, 0.0);
        const std::vector<V3> e = fracDipoles(d);
        for (int i = 0; i < n; i++) {
            const double qi = (qv ? (*qv)[i] : 0.0);
            for (int j1 = 0; j1 < p; j1++) {
                const double t1 = th[0][(size_t) i*p+j1], u1 = dth[0][(size_t) i*p+j1];
                for (int j2 = 0; j2 < p; j2++) {
                    const double t2 = th[1][(size_t) i*p+j2], u2 = dth[1][(size_t) i*p+j2];
                    for (int j3 = 0; j3 < p; j3++) {
                        const double t3 = th[2][(size_t) i*p+j3], u3 = dth[2][(size_t) i*p+j3];
                        Q[gridPoint(i, j1, j2, j3)] +=
                            qi*t1*t2*t3+e[i][0]*u1*t2*t3+e[i][1]*t1*u2*t3+e[i][2]*t1*t2*u3;
                    }
                }
            }
        }
        std::vector<cd> F((size_t) K1*K2*K3h);
        r2c3d(Q.data(), F.data(), K1, K2, K3);
        double energy = 0.0;
        for (int m1 = 0; m1 < K1; m1++)
            for (int m2 = 0; m2 < K2; m2++)
                for (int m3 = 0; m3 < K3h; m3++) {
                    const size_t idx = ((size_t) m1*K2+m2)*K3h+m3;
                    const double g = influence[idx];
                    const double w = (m3 == 0 || (K3%2 == 0 && m3 == K3/2) ? 1.0 : 2.0);
                    const double em = 0.5*w*g*std::norm(F[idx]);
                    energy += em;
                    if (S && g != 0) {   // (core/pgm_core.cpp:222-233)
                        const int f1 = (m1 <= K1/2 ? m1 : m1-K1), f2 = (m2 <= K2/2 ? m2 : m2-K2);
                        V3 mv;
                        for (int k = 0; k < 3; k++)
                            mv[k] = f1*recip[k][0]+f2*recip[k][1]+m3*recip[k][2];
                        const double c2 = 2.0*(kPi*kPi/(s.beta*s.beta)+1.0/dot3(mv, mv));
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++)
                                S[3*a+b] += em*((a == b ? -1.0 : 0.0)+c2*mv[a]*mv[b]);
                    }
                    F[idx] *= g;
                }
        phi.assign((size_t) K1*K2*K3, 0.0);
        c2r3d(F.data(), phi.data(), K1, K2, K3);
        return energy;
    }

    // E -= grad phi at the atoms (core/pgm_core.cpp:241-264)
    void pmeField(const std::vector<double>& phi, std::vector<V3>& E) const {
        const int p = s.order;
        for (int i = 0; i < n; i++) {
            double f[3] = {0, 0, 0};
            for (int j1 = 0; j1 < p; j1++) {
                const double t1 = th[0][(size_t) i*p+j1], u1 = dth[0][(size_t) i*p+j1];
                for (int j2 = 0; j2 < p; j2++) {
                    const double t2 = th[1][(size_t) i*p+j2], u2 = dth[1][(size_t) i*p+j2];
                    for (int j3 = 0; j3 < p; j3++) {
                        const double t3 = th[2][(size_t) i*p+j3], u3 = dth[2][(size_t) i*p+j3];
                        const double ph = phi[gridPoint(i, j1, j2, j3)];
                        f[0] += u1*t2*t3*ph;
                        f[1] += t1*u2*t3*ph;
                        f[2] += t1*t2*u3*ph;
                    }
                }
            }
            for (int k = 0; k < 3; k++)
                E[i][k] -= s.grid[0]*recip[k][0]*f[0]+s.grid[1]*recip[k][1]*f[1]+s.grid[2]*recip[k][2]*f[2];
        }
    }

    // direct + reciprocal + self field of charges qv and dipoles d (core/pgm_core.cpp:268-295)
    std::vector<V3> field(const std::vector<double>& qv, const std::vector<V3>& d) {
        std::vector<V3> E(n, V3{0, 0, 0});
        bool anyQ = false;
        for (double v : qv)
            anyQ = anyQ || v != 0;
        for (const Pair& pr : pairs) {
            const V3& x = pr.x;
            const double qi = qv[pr.i], qj = qv[pr.j];
            const V3& di = d[pr.i];
            const V3& dj = d[pr.j];
            const double djx = dot3(dj, x), dix = dot3(di, x);
            for (int k = 0; k < 3; k++) {
                E[pr.i][k] += qj*x[k]*pr.G[1]-dj[k]*pr.G[1]+djx*x[k]*pr.G[2];
                E[pr.j][k] += -qi*x[k]*pr.G[1]-di[k]*pr.G[1]+dix*x[k]*pr.G[2];
            }
        }
        if (s.periodic) {
            std::vector<double> phi;
            pmePotential(anyQ ? &qv : nullptr, d, phi, nullptr);
            pmeField(phi, E);
            const double cs = 4.0*s.beta*s.beta*s.beta/(3.0*kSqrtPi);
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    E[i][k] += cs*d[i][k];
        }
        return E;
    }

    // The short-range pairs per atom; the preconditioner is on when the radius is set and no atom has more than
    // kPreCap of them (the pairs counted whatever the partners' polarizabilities, core/cuda/pgm_cuda.cu:1139-1154).
    void buildPreconditioner() {
        preOn = false;
        preList.assign(n, {});
        if (!(preRadius > 0))
            return;
        const double r2 = preRadius*preRadius;
        for (size_t m = 0; m < pairs.size(); m++)
            if (dot3(pairs[m].x, pairs[m].x) < r2) {
                preList[pairs[m].i].push_back((int) m);
                preList[pairs[m].j].push_back((int) m);
            }
        preOn = true;
        for (const auto& l : preList)
            if ((int) l.size() > kPreCap)
                preOn = false;
    }

    /* z = M r: z_i = alpha_i (r_i + sum_j T_ij alpha_j r_j) over the short-range partners j with alpha_j > 0,
     * T_ij u = -u G1 + (u.x) x G2 (core/cuda/pgm_cuda.cu:1536-1569); off: z_i = alpha_i r_i, or with the self
     * term alpha_i/(1 - alpha_i c_self) r_i (:1467). */
    void precondition(const std::vector<V3>& r, std::vector<V3>& z) const {
        const double cs = s.periodic ? 4.0*s.beta*s.beta*s.beta/(3.0*kSqrtPi) : 0.0;
        z.assign(n, V3{0, 0, 0});
        for (int i = 0; i < n; i++) {
            const double a = alpha[i];
            if (!preOn) {
                const double f = jacobiSelf ? a/(1.0-a*cs) : a;
                for (int k = 0; k < 3; k++)
                    z[i][k] = f*r[i][k];
                continue;
            }
            V3 e = {0, 0, 0};
            if (a > 0)
                for (int m : preList[i]) {
                    const Pair& pr = pairs[m];
                    const int j = (pr.i == i ? pr.j : pr.i);
                    const double aj = alpha[j];
                    if (aj == 0)
                        continue;
                    const V3 u = {aj*r[j][0], aj*r[j][1], aj*r[j][2]};
                    const double ux = dot3(u, pr.x);
                    for (int k = 0; k < 3; k++)
                        e[k] += -u[k]*pr.G[1]+ux*pr.x[k]*pr.G[2];
                }
            for (int k = 0; k < 3; k++)
                z[i][k] = a*(r[i][k]+e[k]);
        }
    }

    /* PCG with flexible Polak-Ribiere beta on (alpha^-1 - T) mu = E_perm (Jacobi unless a preconditioner is
     * set; the tolerance is on max |alpha r| with every preconditioner, core/cuda/pgm_cuda.cu:1511); predictor,
     * peek and history (core/pgm_core.cpp:297-401). */
    void solve() {
        const std::vector<double> zeroQ(n, 0.0);
        const std::vector<V3> b = field(q, perm);
        double norm = 0;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
                norm += std::fabs(alpha[i]*b[i][k]);
        norm = norm/(3.0*n)+1e-300;

        std::vector<V3> x(n, V3{0, 0, 0});
        const bool samePos = havePos && (historyKey >= 0 ? historyKey == lastKey : lastPos == pos);
        const int h = (int) history.size();
        if (samePos && h > 0)
            x = history[0];
        else if (s.predictor == 2 && h >= 4) {
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    x[i][k] = 4*history[0][i][k]-6*history[1][i][k]+4*history[2][i][k]-history[3][i][k];
        }
        else if (s.predictor >= 1 && h >= 3) {
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    x[i][k] = 3*history[0][i][k]-3*history[1][i][k]+history[2][i][k];
        }
        else if (s.predictor >= 1 && h >= 1)
            x = history[0];
        else
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    x[i][k] = alpha[i]*b[i][k];

        auto applyA = [&](const std::vector<V3>& v) {
            const std::vector<V3> Tv = field(zeroQ, v);
            std::vector<V3> out(n);
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    out[i][k] = (alpha[i] > 0 ? v[i][k]/alpha[i]-Tv[i][k] : 0.0);
            return out;
        };
        auto vdot = [&](const std::vector<V3>& u, const std::vector<V3>& v) {
            double acc = 0;
            for (int i = 0; i < n; i++)
                acc += dot3(u[i], v[i]);
            return acc;
        };
        auto errOf = [&](const std::vector<V3>& r) {
            double m = 0;
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    m = std::max(m, std::fabs(alpha[i]*r[i][k]));
            return m/norm;
        };
        buildPreconditioner();
        const std::vector<V3> Ax = applyA(x);
        std::vector<V3> r(n), z(n), dir;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
                r[i][k] = (alpha[i] > 0 ? b[i][k]-Ax[i][k] : 0.0);
        precondition(r, z);
        dir = z;
        double rz = vdot(r, z), err = errOf(r);
        int it = 0;
        while (err > s.tol && it < s.maxIter) {
            const std::vector<V3> Ap = applyA(dir);
            const double step = rz/std::max(vdot(dir, Ap), 1e-300);
            std::vector<V3> rn(n), zn;
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++) {
                    x[i][k] += step*dir[i][k];
                    rn[i][k] = r[i][k]-step*Ap[i][k];
                }
            precondition(rn, zn);
            double bnum = 0;
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    bnum += zn[i][k]*(rn[i][k]-r[i][k]);
            const double bcoef = bnum/std::max(rz, 1e-300);
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    dir[i][k] = zn[i][k]+bcoef*dir[i][k];
            r.swap(rn);
            z.swap(zn);
            rz = vdot(r, z);
            err = errOf(r);
            it++;
        }
        if (s.peek != 0)
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    x[i][k] += s.peek*alpha[i]*r[i][k];
        mu = x;
        lastIterations = it;
        lastResidual = err;
        if (!samePos) {
            history.push_front(mu);
            while (history.size() > 4)
                history.pop_back();
        }
        lastPos = pos;
        lastKey = historyKey;
        havePos = true;
    }

    // energy terms, forces, strain derivative at the solved dipoles (core/pgm_core.cpp:405-537)
    double compute(const double* p, const double* box9, double* forcesAdd) {
        setGeometry(p, box9);
        if (s.polarizable)
            solve();
        else {
            mu.assign(n, V3{0, 0, 0});
            lastIterations = 0;
        }
        std::vector<V3> d(n);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
                d[i][k] = perm[i][k]+mu[i][k];
        double T[5] = {0, 0, 0, 0, 0};
        std::vector<V3> E(n, V3{0, 0, 0}), F(n, V3{0, 0, 0});
        double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (const Pair& pr : pairs) {
            const V3& x = pr.x;
            const double qi = q[pr.i], qj = q[pr.j];
            const V3& di = d[pr.i];
            const V3& dj = d[pr.j];
            const double dix = dot3(di, x), djx = dot3(dj, x), didj = dot3(di, dj);
            const double* G = pr.G;
            T[0] += qi*qj*G[0]+(qi*djx-qj*dix)*G[1]+didj*G[1]-dix*djx*G[2];
            const double radial = -qi*qj*G[1]-(qi*djx-qj*dix)*G[2]-didj*G[2]+dix*djx*G[3];
            for (int k = 0; k < 3; k++) {
                const double g = radial*x[k]+(qi*dj[k]-qj*di[k])*G[1]-(djx*di[k]+dix*dj[k])*G[2];
                for (int b = 0; b < 3; b++)
                    S[3*k+b] += g*x[b];
                F[pr.i][k] -= g;
                F[pr.j][k] += g;
                E[pr.i][k] += qj*x[k]*G[1]-dj[k]*G[1]+djx*x[k]*G[2];
                E[pr.j][k] += -qi*x[k]*G[1]-di[k]*G[1]+dix*x[k]*G[2];
            }
        }
        if (s.periodic) {
            std::vector<double> phi;
            T[1] = pmePotential(&q, d, phi, S);
            std::vector<V3> Erec(n, V3{0, 0, 0});
            pmeField(phi, Erec);
            for (int i = 0; i < n; i++)
                for (int a = 0; a < 3; a++) {
                    E[i][a] += Erec[i][a];
                    for (int b = 0; b < 3; b++)
                        S[3*a+b] += Erec[i][a]*d[i][b];
                }
            // reciprocal forces -[q grad phi + (grad grad phi) . d] (core/pgm_core.cpp:452-487)
            const int p = s.order;
            const std::vector<V3> e = fracDipoles(d);
            for (int i = 0; i < n; i++) {
                double g1[3] = {0, 0, 0}, g2[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int j1 = 0; j1 < p; j1++) {
                    const size_t o1 = (size_t) i*p+j1;
                    const double t1 = th[0][o1], u1 = dth[0][o1], v1 = d2th[0][o1];
                    for (int j2 = 0; j2 < p; j2++) {
                        const size_t o2 = (size_t) i*p+j2;
                        const double t2 = th[1][o2], u2 = dth[1][o2], v2 = d2th[1][o2];
                        for (int j3 = 0; j3 < p; j3++) {
                            const size_t o3 = (size_t) i*p+j3;
                            const double t3 = th[2][o3], u3 = dth[2][o3], v3 = d2th[2][o3];
                            const double ph = phi[gridPoint(i, j1, j2, j3)];
                            g1[0] += u1*t2*t3*ph;
                            g1[1] += t1*u2*t3*ph;
                            g1[2] += t1*t2*u3*ph;
                            g2[0][0] += v1*t2*t3*ph;
                            g2[1][1] += t1*v2*t3*ph;
                            g2[2][2] += t1*t2*v3*ph;
                            g2[0][1] += u1*u2*t3*ph;
                            g2[0][2] += u1*t2*u3*ph;
                            g2[1][2] += t1*u2*u3*ph;
                        }
                    }
                }
                g2[1][0] = g2[0][1];
                g2[2][0] = g2[0][2];
                g2[2][1] = g2[1][2];
                for (int a = 0; a < 3; a++) {
                    const double c = q[i]*g1[a]+e[i][0]*g2[a][0]+e[i][1]*g2[a][1]+e[i][2]*g2[a][2];
                    for (int k = 0; k < 3; k++)
                        F[i][k] -= s.grid[a]*recip[k][a]*c;
                }
            }
            // self and neutralizing background (core/pgm_core.cpp:488-500)
            const double cs = 4.0*s.beta*s.beta*s.beta/(3.0*kSqrtPi);
            double q2 = 0, d2 = 0, Qt = 0;
            for (int i = 0; i < n; i++) {
                q2 += q[i]*q[i];
                d2 += dot3(d[i], d[i]);
                Qt += q[i];
                for (int k = 0; k < 3; k++)
                    E[i][k] += cs*d[i][k];
            }
            T[2] = -(s.beta/kSqrtPi)*q2-0.5*cs*d2;
            T[3] = -kPi*Qt*Qt/(2.0*volume*s.beta*s.beta);
            for (int a = 0; a < 3; a++)
                S[4*a] -= T[3];
        }
        for (int i = 0; i < n; i++)
            if (alpha[i] > 0)
                T[4] += dot3(mu[i], mu[i])/(2.0*alpha[i]);
        for (int k = 0; k < 5; k++)
            terms[k] = s.ke*T[k];
        if (forcesAdd) {
            // covalent chain rule: dE/dr_partner = c (g - (g.u) u)/|v|, g = -E_atom (core/pgm_core.cpp:512-529)
            for (size_t m = 0; m < covAtom.size(); m++) {
                const int i = covAtom[m], j = covPartner[m];
                V3 v = {pos[j][0]-pos[i][0], pos[j][1]-pos[i][1], pos[j][2]-pos[i][2]};
                if (s.periodic)
                    minImage(v, box);
                const double len = std::sqrt(dot3(v, v));
                const V3 u = {v[0]/len, v[1]/len, v[2]/len};
                const V3 g = {-E[i][0], -E[i][1], -E[i][2]};
                const double gu = dot3(g, u);
                for (int k = 0; k < 3; k++) {
                    const double t = covStrength[m]*(g[k]-gu*u[k])/len;
                    F[j][k] -= t;
                    F[i][k] += t;
                    for (int b = 0; b < 3; b++)
                        S[3*k+b] += t*v[b];
                }
            }
            for (int k = 0; k < 9; k++)
                strain[k] = s.ke*S[k];
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    forcesAdd[3*i+k] += s.ke*F[i][k];
        }
        return terms[0]+terms[1]+terms[2]+terms[3]+terms[4];
    }
};

Engine* E(void* h) {
    if (!h)
        throw std::invalid_argument("pgm_oracle: null handle");
    return static_cast<Engine*>(h);
}

template<class Fn>
int guarded(Fn fn) {
    try {
        fn();
        return 0;
    }
    catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

}   // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
const char* orc_kind(void) { return "port"; }

void* orc_create(const orc_settings* s, int n, const double* q, const double* radius, const double* alpha, int ncov,
                 const int* covAtom, const int* covPartner, const double* covStrength) {
    Engine* e = nullptr;
    int rc = guarded([&] {
        if (!s || n <= 0 || !q || !radius || !alpha)
            throw std::invalid_argument("pgm_oracle: bad arguments");
        if (s->periodic) {   // core/pgm_core.cpp:30-38
            for (int a = 0; a < 3; a++)
                if (s->grid[a] <= 0)
                    throw std::invalid_argument("pgm_oracle: PME grid not set");
            if (s->order < 3)
                throw std::invalid_argument("pgm_oracle: PME order must be >= 3");
            if (s->beta <= 0)
                throw std::invalid_argument("pgm_oracle: Ewald coefficient not set");
        }
        e = new Engine();
        e->s = *s;
        e->n = n;
        e->setTopology(q, radius, alpha, ncov, covAtom, covPartner, covStrength);
        e->perm.assign(n, V3{0, 0, 0});
        e->mu.assign(n, V3{0, 0, 0});
    });
    if (rc != 0) {
        delete e;
        return nullptr;
    }
    return e;
}

void orc_destroy(void* h) { delete static_cast<Engine*>(h); }

int orc_compute(void* h, const double* pos, const double* box9, double* forcesAdd, double* energy) {
    return guarded([&] {
        const double en = E(h)->compute(pos, box9, forcesAdd);
        if (energy)
            *energy = en;
    });
}

int orc_terms(void* h, double* out5) {
    return guarded([&] { std::memcpy(out5, E(h)->terms, 5*sizeof(double)); });
}

int orc_induced(void* h, double* mu) {
    return guarded([&] {
        Engine* e = E(h);
        for (int i = 0; i < e->n; i++)
            for (int k = 0; k < 3; k++)
                mu[3*i+k] = e->mu[i][k];
    });
}

int orc_permanent(void* h, double* p) {
    return guarded([&] {
        Engine* e = E(h);
        for (int i = 0; i < e->n; i++)
            for (int k = 0; k < 3; k++)
                p[3*i+k] = e->perm[i][k];
    });
}

int orc_iterations(void* h) { return h ? static_cast<Engine*>(h)->lastIterations : -1; }
double orc_residual(void* h) { return h ? static_cast<Engine*>(h)->lastResidual : -1.0; }

int orc_strain(void* h, double* out9) {
    return guarded([&] { std::memcpy(out9, E(h)->strain, 9*sizeof(double)); });
}

int orc_set_history_key(void* h, long long key) {
    return guarded([&] { E(h)->historyKey = key; });
}

int orc_reset_history(void* h) {
    return guarded([&] {
        E(h)->history.clear();
        E(h)->havePos = false;
    });
}

int orc_push_history(void* h, const double* mu) {
    return guarded([&] {
        Engine* e = E(h);
        std::vector<V3> m(e->n);
        for (int i = 0; i < e->n; i++)
            m[i] = V3{mu[3*i], mu[3*i+1], mu[3*i+2]};
        e->history.push_front(m);
        while (e->history.size() > 4)
            e->history.pop_back();
    });
}

int orc_set_topology(void* h, const double* q, const double* radius, const double* alpha, int ncov,
                     const int* covAtom, const int* covPartner, const double* covStrength) {
    return guarded([&] {
        E(h)->setTopology(q, radius, alpha, ncov, covAtom, covPartner, covStrength);
        E(h)->havePos = false;   // core/pgm_core.cpp:41-47
    });
}

int orc_set_geometry(void* h, const double* pos, const double* box9) {
    return guarded([&] { E(h)->setGeometry(pos, box9); });
}

int orc_field(void* h, const double* q, const double* d, double* out) {
    return guarded([&] {
        Engine* e = E(h);
        if (e->pos.empty())
            throw std::logic_error("pgm_oracle: field before set_geometry");
        std::vector<double> qv(q, q+e->n);
        std::vector<V3> dv(e->n);
        for (int i = 0; i < e->n; i++)
            dv[i] = V3{d[3*i], d[3*i+1], d[3*i+2]};
        const std::vector<V3> f = e->field(qv, dv);
        for (int i = 0; i < e->n; i++)
            for (int k = 0; k < 3; k++)
                out[3*i+k] = f[i][k];
    });
}

long long orc_num_pairs(void* h) { return h ? (long long) static_cast<Engine*>(h)->pairs.size() : -1; }

int orc_set_preconditioner(void* h, double radius, int jacobiSelf) {
    return guarded([&] {
        Engine* e = E(h);
        if (!(radius >= 0) || (e->s.periodic && radius > e->s.cutoff))
            throw std::invalid_argument("pgm_oracle: the preconditioner radius must be in [0, cutoff]");
        e->preRadius = radius;
        e->jacobiSelf = jacobiSelf != 0;
    });
}

int orc_precondition(void* h, const double* r, double* z) {
    return guarded([&] {
        Engine* e = E(h);
        if (e->pos.empty())
            throw std::logic_error("pgm_oracle: precondition before set_geometry");
        e->buildPreconditioner();
        std::vector<V3> rv(e->n), zv;
        for (int i = 0; i < e->n; i++)
            rv[i] = V3{r[3*i], r[3*i+1], r[3*i+2]};
        e->precondition(rv, zv);
        for (int i = 0; i < e->n; i++)
            for (int k = 0; k < 3; k++)
                z[3*i+k] = zv[i][k];
    });
}

int orc_preconditioner_on(void* h) { return h ? (int) static_cast<Engine*>(h)->preOn : -1; }

int orc_erf_kernels(double a, double r, int nmax, double* B) {
    return guarded([&] {
        if (nmax < 1 || nmax > 4)
            throw std::invalid_argument("pgm_oracle: nmax must be 1..4");
        erfKernels(a, r, nmax, B);
    });
}

int orc_erfc_kernels(double a, double r, int nmax, double* B) {
    return guarded([&] {
        if (nmax < 1 || nmax > 4)
            throw std::invalid_argument("pgm_oracle: nmax must be 1..4");
        erfcKernels(a, r, nmax, B);
    });
}

int orc_pair_kernels(double a, double beta, double r, int nmax, double* G) {
    return guarded([&] {
        if (nmax < 1 || nmax > 4)
            throw std::invalid_argument("pgm_oracle: nmax must be 1..4");
        pairKernels(a, beta, r, nmax, G);
    });
}

int orc_gvdw_g(double y, double* G, double* H) {
    return guarded([&] { gvdwG(y, *G, *H); });
}

int orc_gvdw_pair(double r, double bij, double A, double C6, double B, int rep, double* energy, double* dEdR) {
    return guarded([&] { *energy = gvdwPair(r, bij, A, C6, B, rep, *dEdR); });
}

int orc_gvdw_sum(int n, const double* pos, const double* box9, const double* radius, const double* sqrtA,
                 const double* sqrtC6, const double* b, int rep, double rvdw, int nskip, const int* skipI,
                 const int* skipJ, double* energy, double* forcesAdd, double* strain9) {
    return guarded([&] {
        double box[3][3];
        std::memcpy(box, box9, 9*sizeof(double));
        std::vector<std::vector<int>> skip(n);
        for (int k = 0; k < nskip; k++) {
            if (skipI[k] < 0 || skipI[k] >= n || skipJ[k] < 0 || skipJ[k] >= n || skipI[k] == skipJ[k])
                throw std::invalid_argument("pgm_oracle: skipped pair out of range");
            skip[std::min(skipI[k], skipJ[k])].push_back(std::max(skipI[k], skipJ[k]));
        }
        double E = 0, S[9] = {0};
        for (int i = 0; i < n; i++)
            for (int j = i+1; j < n; j++) {
                const double A = sqrtA[i]*sqrtA[j], C6 = sqrtC6[i]*sqrtC6[j];
                if (A == 0 && C6 == 0)
                    continue;
                if (std::find(skip[i].begin(), skip[i].end(), j) != skip[i].end())
                    continue;
                V3 x = {pos[3*i]-pos[3*j], pos[3*i+1]-pos[3*j+1], pos[3*i+2]-pos[3*j+2]};
                minImage(x, box);
                const double r = std::sqrt(dot3(x, x));
                if (r >= rvdw)
                    continue;
                const double bij = 1.0/std::sqrt(2.0*(radius[i]*radius[i]+radius[j]*radius[j]));
                double dEdR;
                E += gvdwPair(r, bij, A, C6, 0.5*(b[i]+b[j]), rep, dEdR);
                for (int a = 0; a < 3; a++) {
                    if (forcesAdd) {
                        forcesAdd[3*i+a] += dEdR*x[a];
                        forcesAdd[3*j+a] -= dEdR*x[a];
                    }
                    for (int c = 0; c < 3; c++)
                        S[3*a+c] -= dEdR*x[a]*x[c];   // dU/d eps_ac = (dU/dr) x_a x_c / r
                }
            }
        if (energy)
            *energy = E;
        if (strain9)
            for (int k = 0; k < 9; k++)
                strain9[k] += S[k];
    });
}

int orc_bspline_moduli(int K, int p, double* out) {
    return guarded([&] {
        const std::vector<double> m = bsplineModuli(K, p);
        std::copy(m.begin(), m.end(), out);
    });
}

double orc_ewald_coefficient(double rc, double rtol) { return ewaldCoefficient(rc, rtol); }

int orc_fft3d_r2c(int K1, int K2, int K3, const double* in, double* out) {
    return guarded([&] { r2c3d(in, reinterpret_cast<cd*>(out), K1, K2, K3); });
}

int orc_fft3d_c2r(int K1, int K2, int K3, const double* in, double* out) {
    return guarded([&] { c2r3d(reinterpret_cast<const cd*>(in), out, K1, K2, K3); });
}

}   // extern "C"
