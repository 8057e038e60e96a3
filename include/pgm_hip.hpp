/* pgm_hip.hpp: header-only C++ facade over the C ABI of pgm_hip.h, with the reference's two C++ APIs:
 *
 *   pgmhip::Engine        the method set of pgmcuda::Engine (core/cuda/pgm_cuda.h:47-127)
 *   pgmhip::core::Engine  the method set of pgmcore::Engine (core/pgm_core.h:47-115), periodic (PME) only
 *
 * so code written against either reference API switches by changing the namespace and linking
 * libpgm_hip.so.  Errors are rethrown as std::runtime_error with the engine's message ("pgm_hip: ...").
 */
#ifndef PGM_HIP_HPP_
#define PGM_HIP_HPP_

#include "pgm_hip.h"

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

namespace pgmhip {

/* pgmcuda::Settings (core/cuda/pgm_cuda.h:25-39) plus the precision and polarizable switches. */
struct Settings {
    double cutoff = 0.9;
    double beta = 3.5;
    int grid[3] = {36, 36, 36};
    int order = 6;
    double tol = 1e-5;
    int maxIter = 100;
    double peek = 0.65;
    int predictor = 2;
    double ke = 138.935456;
    int device = 0;
    double skin = 0.15;
    bool mixed = true;
    bool polarizable = true;
};

/* pgmcuda::Terms (core/cuda/pgm_cuda.h:41-45). */
struct Terms {
    double direct = 0, reciprocal = 0, self = 0, background = 0, polarization = 0, gvdw = 0;
    double total() const { return direct+reciprocal+self+background+polarization+gvdw; }
};

namespace detail {
inline pgm_settings toC(const Settings& s) {
    pgm_settings c;
    pgm_default_settings(&c);
    c.cutoff = s.cutoff;
    c.beta = s.beta;
    for (int a = 0; a < 3; a++)
        c.grid[a] = s.grid[a];
    c.order = s.order;
    c.tol = s.tol;
    c.max_iter = s.maxIter;
    c.peek = s.peek;
    c.predictor = s.predictor;
    c.ke = s.ke;
    c.device = s.device;
    c.skin = s.skin;
    c.precision = s.mixed ? PGM_MIXED : PGM_DOUBLE;
    c.polarizable = s.polarizable ? 1 : 0;
    return c;
}
inline Terms fromC(const pgm_terms& t) {
    Terms r;
    r.direct = t.direct;
    r.reciprocal = t.reciprocal;
    r.self = t.self;
    r.background = t.background;
    r.polarization = t.polarization;
    r.gvdw = t.gvdw;
    return r;
}
}   // namespace detail

/* pgmcuda::Engine: the caller provides atoms and covalent dipoles; the engine keeps its own pair list. */
class Engine {
public:
    explicit Engine(const Settings& s) {
        const pgm_settings c = detail::toC(s);
        e_ = pgm_create(&c);
        if (!e_)
            throw std::runtime_error(pgm_last_error(nullptr));
    }
    ~Engine() { pgm_destroy(e_); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void setAtoms(int n, const double* q, const double* radius, const double* alpha, const int* globalIndex,
                  int numGlobal) {
        check(pgm_set_atoms(e_, n, q, radius, alpha, globalIndex, numGlobal));
        n_ = n;
    }
    void setGvdw(int repulsion, double rvdw, const double* sqrtA, const double* sqrtC6, const double* b, int nskip,
                 const int* skipI, const int* skipJ) {
        check(pgm_set_gvdw(e_, repulsion, rvdw, sqrtA, sqrtC6, b, nskip, skipI, skipJ));
    }
    void setCovalent(int ncov, const int* atom, const int* partner, const double* strength) {
        check(pgm_set_covalent(e_, ncov, atom, partner, strength));
    }
    void setCandidates(int nseg, const int* segI, const int* segShift, const int* segBegin, const int* candJ) {
        check(pgm_set_candidates(e_, nseg, segI, segShift, segBegin, candJ));
    }
    Terms compute(const double* pos, const double box[3][3], bool newStep, double* f, double* strain) {
        pgm_terms t;
        check(pgm_compute(e_, pos, &box[0][0], newStep ? 1 : 0, f, strain, &t));
        return detail::fromC(t);
    }
    Terms computeF(const float* pos, const double box[3][3], bool newStep, double* strain) {
        pgm_terms t;
        check(pgm_compute_f32(e_, pos, &box[0][0], newStep ? 1 : 0, strain, &t));
        forces_.resize(3*(size_t) n_);
        check(pgm_forces(e_, forces_.data()));
        return detail::fromC(t);
    }
    const double* forces() const { return forces_.data(); }
    Terms computeDevice(const void* posq4, const double box[3][3], bool newStep, long long* forceFixed, int stride,
                        double fixedScale, void* stream) {
        pgm_terms t;
        check(pgm_compute_device(e_, posq4, &box[0][0], newStep ? 1 : 0, forceFixed, stride, fixedScale, stream, &t));
        return detail::fromC(t);
    }
    void launchDevice(const void* posq4, const double box[3][3], bool newStep, long long* forceFixed, int stride,
                      double fixedScale, void* stream) {
        check(pgm_launch_device(e_, posq4, &box[0][0], newStep ? 1 : 0, forceFixed, stride, fixedScale, stream));
    }
    bool finishDevice(Terms& terms) {
        pgm_terms t;
        int valid = 0;
        check(pgm_finish_device(e_, &t, &valid));
        terms = detail::fromC(t);
        return valid != 0;
    }
    void waitDevice(void* stream) { check(pgm_wait_device(e_, stream)); }
    // computeF copies the caller's array (nothing is registered): nothing to release.
    void releasePositions() {}
    int iterations() const { return pgm_iterations(e_); }
    double residual() const { return pgm_residual(e_); }
    void inducedDipoles(double* mu) const { check(pgm_induced_dipoles(e_, mu)); }
    void permanentDipoles(double* p) const { check(pgm_permanent_dipoles(e_, p)); }
    void resetHistory() { check(pgm_reset_history(e_)); }
    /* No counterpart in pgmcuda::Engine (pgmcore::Engine::field, core/pgm_core.h:76-77): the field of charges q and
     * dipoles d at every atom at positions pos (pgm_field). */
    void field(const double* pos, const double box[3][3], const double* q, const double* d, double* out) {
        check(pgm_field(e_, pos, &box[0][0], q, d, out));
    }
    long listBuilds() const { return pgm_list_builds(e_); }
    int numCandidates() const { return (int) pgm_num_candidates(e_); }
    long lateListChecks() const { return pgm_late_list_checks(e_); }
    /* No counterpart in the reference: parameter gradients of a loss L(E, F) (pgm_backward). */
    void backward(const double* pos, const double box[3][3], const double* dLdF, double dLdE, double* gradQ,
                  double* gradRadius, double* gradAlpha, double* gradCov) {
        check(pgm_backward(e_, pos, &box[0][0], dLdF, dLdE, gradQ, gradRadius, gradAlpha, gradCov));
    }
    pgm_engine* handle() const { return e_; }

private:
    pgm_engine* e_ = nullptr;
    int n_ = 0;
    std::vector<double> forces_;
    void check(int rc) const {
        if (rc != 0)
            throw std::runtime_error(pgm_last_error(e_));
    }
};

namespace core {

typedef std::array<double, 3> V3;

/* pgmcore::Settings (core/pgm_core.h:22-34); periodic must be true. */
struct Settings {
    bool periodic = true;
    double cutoff = 0.9;
    double beta = 0.0;
    int grid[3] = {0, 0, 0};
    int order = 6;
    bool polarizable = true;
    double tol = 1e-5;
    int maxIter = 100;
    double peek = 0.65;
    int predictor = 2;
    double ke = 138.935456;
    bool mixed = false;   // FP64 by default, as pgmcore
    int device = 0;
};

/* pgmcore::Topology (core/pgm_core.h:36-40). */
struct Topology {
    std::vector<double> q, radius, alpha;
    std::vector<int> covAtom, covPartner;
    std::vector<double> covStrength;
};

/* pgmcore::EnergyTerms (core/pgm_core.h:42-45). */
struct EnergyTerms {
    double direct = 0, reciprocal = 0, self = 0, background = 0, polarization = 0;
    double total() const { return direct+reciprocal+self+background+polarization; }
};

/* pgmcore::Engine on the GPU: compute() adds forces, one predictor entry per history key. */
class Engine {
public:
    Engine(const Topology& top, const Settings& s) : s_(s) {
        if (!s.periodic)
            throw std::invalid_argument("pgm_hip: only periodic (PME) systems run on the GPU");
        pgmhip::Settings g;
        g.cutoff = s.cutoff;
        g.beta = s.beta;
        for (int a = 0; a < 3; a++)
            g.grid[a] = s.grid[a];
        g.order = s.order;
        g.polarizable = s.polarizable;
        g.tol = s.tol;
        g.maxIter = s.maxIter;
        g.peek = s.peek;
        g.predictor = s.predictor;
        g.ke = s.ke;
        g.mixed = s.mixed;
        g.device = s.device;
        eng_ = new pgmhip::Engine(g);
        setTopology(top);
    }
    ~Engine() { delete eng_; }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    /* Energy [kJ/mol]; forces added to *forces when given, and then the strain derivative too
     * (core/pgm_core.h:47-51, 62-71). */
    double compute(const std::vector<V3>& pos, const V3 box[3], std::vector<V3>* forces) {
        const int n = (int) top_.q.size();
        if ((int) pos.size() != n)
            throw std::invalid_argument("pgm_hip: positions do not match the topology");
        double b[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                b[r][c] = box[r][c];
        // same key, or no key and unchanged positions: the same step (core/pgm_core.h:56-61)
        const bool same = havePos_ && (key_ >= 0 ? key_ == lastKey_ : pos == lastPos_);
        std::vector<double> f(forces ? 3*(size_t) n : 0, 0.0);
        std::array<double, 9> S{};
        const Terms t = eng_->compute(&pos[0][0], b, !same, forces ? f.data() : nullptr, forces ? S.data() : nullptr);
        if (forces) {
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    (*forces)[i][k] += f[3*i+k];
            lastStrain_ = S;
        }
        terms_.direct = t.direct;
        terms_.reciprocal = t.reciprocal;
        terms_.self = t.self;
        terms_.background = t.background;
        terms_.polarization = t.polarization;
        lastPos_ = pos;
        lastKey_ = key_;
        havePos_ = true;
        setGeometry(pos, box);
        return terms_.total();
    }
    void setHistoryKey(long long key) { key_ = key; }
    /* Field [e/nm^2] of charges q and dipoles d at every atom (direct + reciprocal + self), at the positions of
     * the last setGeometry or compute (core/pgm_core.h:76-79). */
    std::vector<V3> field(const std::vector<double>& q, const std::vector<V3>& d) {
        const int n = (int) top_.q.size();
        if (!haveGeometry_)
            throw std::logic_error("pgm_hip: field before setGeometry or compute");
        if ((int) q.size() != n || (int) d.size() != n)
            throw std::invalid_argument("pgm_hip: field operands do not match the topology");
        std::vector<V3> out(n);
        eng_->field(&geoPos_[0][0], geoBox_, q.data(), &d[0][0], &out[0][0]);
        return out;
    }
    /* The positions and box of the next field() (core/pgm_core.h:78-79); compute() sets them too. */
    void setGeometry(const std::vector<V3>& pos, const V3 box[3]) {
        if (pos.size() != top_.q.size())
            throw std::invalid_argument("pgm_hip: positions do not match the topology");
        geoPos_ = pos;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                geoBox_[r][c] = box[r][c];
        haveGeometry_ = true;
    }
    const std::vector<V3>& permanentDipoles() {
        perm_.resize(top_.q.size());
        eng_->permanentDipoles(&perm_[0][0]);
        return perm_;
    }
    const std::vector<V3>& inducedDipoles() {
        mu_.resize(top_.q.size());
        eng_->inducedDipoles(&mu_[0][0]);
        return mu_;
    }
    int iterations() const { return eng_->iterations(); }
    double residual() const { return eng_->residual(); }
    const EnergyTerms& terms() const { return terms_; }
    /* Strain derivative of the last compute(), row-major (GROMACS virial = S^T/2). */
    const std::array<double, 9>& strainDerivative() const { return lastStrain_; }
    void setTopology(const Topology& top) {
        top_ = top;
        const int n = (int) top.q.size();
        eng_->setAtoms(n, top.q.data(), top.radius.data(), top.alpha.data(), nullptr, n);
        eng_->setCovalent((int) top.covAtom.size(), top.covAtom.data(), top.covPartner.data(),
                          top.covStrength.data());
        havePos_ = false;
    }
    void resetHistory() {
        eng_->resetHistory();
        havePos_ = false;
    }
private:
    Settings s_;
    Topology top_;
    pgmhip::Engine* eng_ = nullptr;
    std::vector<V3> perm_, mu_, lastPos_;
    long long key_ = -1, lastKey_ = -1;
    bool havePos_ = false;
    EnergyTerms terms_;
    std::array<double, 9> lastStrain_{};
    std::vector<V3> geoPos_;
    double geoBox_[3][3] = {{0}};
    bool haveGeometry_ = false;
};

}   // namespace core

}   // namespace pgmhip

#endif
