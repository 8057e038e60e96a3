/* ref_driver: test infrastructure only.
 *
 * Exports the C ABI of pgm_oracle.h over the reference's own FP64 engine pgmcore::Engine
 * (core/pgm_core.h:47-115), its numerics (core/pgm_math.h) and its FFT (core/pocketfft_hdronly.h), so the
 * tests can check the restatement in pgm_oracle.cpp against the reference itself.  Built by
 * oracle/Makefile together with core/pgm_core.cpp, compiled where it lies; the output goes to oracle/_ref/.
 */
#include "pgm_oracle.h"

#include "gvdw_kernel.h"
#include "pgm_core.h"
#include "pgm_math.h"
#include "pocketfft_hdronly.h"

#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct Handle {
    pgmcore::Topology top;
    pgmcore::Settings s;
    std::unique_ptr<pgmcore::Engine> eng;
    int n = 0;
    bool haveGeometry = false;
};

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

Handle* H(void* h) {
    if (!h)
        throw std::invalid_argument("ref_driver: null handle");
    return static_cast<Handle*>(h);
}

void fillTop(pgmcore::Topology& t, int n, const double* q, const double* r, const double* a, int ncov, const int* ca,
             const int* cp, const double* cs) {
    t.q.assign(q, q+n);
    t.radius.assign(r, r+n);
    t.alpha.assign(a, a+n);
    t.covAtom.assign(ca, ca+ncov);
    t.covPartner.assign(cp, cp+ncov);
    t.covStrength.assign(cs, cs+ncov);
}

std::vector<pgmcore::V3> toV3(const double* p, int n) {
    std::vector<pgmcore::V3> v(n);
    for (int i = 0; i < n; i++)
        v[i] = {p[3*i], p[3*i+1], p[3*i+2]};
    return v;
}

void toBox(const double* b9, pgmcore::V3 box[3]) {
    for (int r = 0; r < 3; r++)
        box[r] = {b9[3*r], b9[3*r+1], b9[3*r+2]};
}

void copyOut(const std::vector<pgmcore::V3>& v, double* out) {
    for (size_t i = 0; i < v.size(); i++)
        for (int k = 0; k < 3; k++)
            out[3*i+k] = v[i][k];
}

}   // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
const char* orc_kind(void) { return "reference"; }

void* orc_create(const orc_settings* s, int n, const double* q, const double* radius, const double* alpha, int ncov,
                 const int* covAtom, const int* covPartner, const double* covStrength) {
    Handle* h = nullptr;
    int rc = guarded([&] {
        h = new Handle();
        h->n = n;
        h->s.periodic = s->periodic != 0;
        h->s.cutoff = s->cutoff;
        h->s.beta = s->beta;
        for (int a = 0; a < 3; a++)
            h->s.grid[a] = s->grid[a];
        h->s.order = s->order;
        h->s.polarizable = s->polarizable != 0;
        h->s.tol = s->tol;
        h->s.maxIter = s->maxIter;
        h->s.peek = s->peek;
        h->s.predictor = s->predictor;
        h->s.ke = s->ke;
        fillTop(h->top, n, q, radius, alpha, ncov, covAtom, covPartner, covStrength);
        h->eng.reset(new pgmcore::Engine(h->top, h->s));
    });
    if (rc != 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

int orc_compute(void* h, const double* pos, const double* box9, double* forcesAdd, double* energy) {
    return guarded([&] {
        Handle* e = H(h);
        pgmcore::V3 box[3] = {};
        if (box9)
            toBox(box9, box);
        std::vector<pgmcore::V3> f(e->n, pgmcore::V3{0, 0, 0});
        const double en = e->eng->compute(toV3(pos, e->n), box, forcesAdd ? &f : nullptr);
        if (forcesAdd)
            for (int i = 0; i < e->n; i++)
                for (int k = 0; k < 3; k++)
                    forcesAdd[3*i+k] += f[i][k];
        if (energy)
            *energy = en;
        e->haveGeometry = true;
    });
}

int orc_terms(void* h, double* out5) {
    return guarded([&] {
        const pgmcore::EnergyTerms& t = H(h)->eng->terms();
        out5[0] = t.direct;
        out5[1] = t.reciprocal;
        out5[2] = t.self;
        out5[3] = t.background;
        out5[4] = t.polarization;
    });
}

int orc_induced(void* h, double* mu) {
    return guarded([&] { copyOut(H(h)->eng->inducedDipoles(), mu); });
}

int orc_permanent(void* h, double* p) {
    return guarded([&] { copyOut(H(h)->eng->permanentDipoles(), p); });
}

int orc_iterations(void* h) { return h ? static_cast<Handle*>(h)->eng->iterations() : -1; }
double orc_residual(void* h) { return h ? static_cast<Handle*>(h)->eng->residual() : -1.0; }

int orc_strain(void* h, double* out9) {
    return guarded([&] {
        const std::array<double, 9>& S = H(h)->eng->strainDerivative();
        std::memcpy(out9, S.data(), 9*sizeof(double));
    });
}

int orc_set_history_key(void* h, long long key) {
    return guarded([&] { H(h)->eng->setHistoryKey(key); });
}

int orc_reset_history(void* h) {
    return guarded([&] { H(h)->eng->resetHistory(); });
}

int orc_push_history(void*, const double*) {
    g_err = "ref_driver: pgmcore::Engine has no history import";
    return -1;
}

int orc_set_topology(void* h, const double* q, const double* radius, const double* alpha, int ncov,
                     const int* covAtom, const int* covPartner, const double* covStrength) {
    return guarded([&] {
        Handle* e = H(h);
        fillTop(e->top, e->n, q, radius, alpha, ncov, covAtom, covPartner, covStrength);
        e->eng->setTopology(e->top);
    });
}

int orc_set_geometry(void* h, const double* pos, const double* box9) {
    return guarded([&] {
        Handle* e = H(h);
        pgmcore::V3 box[3] = {};
        if (box9)
            toBox(box9, box);
        e->eng->setGeometry(toV3(pos, e->n), box);
        e->haveGeometry = true;
    });
}

int orc_field(void* h, const double* q, const double* d, double* out) {
    return guarded([&] {
        Handle* e = H(h);
        if (!e->haveGeometry)
            throw std::logic_error("ref_driver: field before set_geometry");
        std::vector<double> qv(q, q+e->n);
        copyOut(e->eng->field(qv, toV3(d, e->n)), out);
    });
}

long long orc_num_pairs(void*) { return -1; }   // pgmcore::Engine keeps its pairs private

// The preconditioners are options of the reference's CUDA engine only; pgmcore::Engine is Jacobi.
int orc_set_preconditioner(void*, double, int) {
    g_err = "pgm_ref: pgmcore::Engine has no preconditioner option";
    return -1;
}
int orc_precondition(void*, const double*, double*) {
    g_err = "pgm_ref: pgmcore::Engine has no preconditioner option";
    return -1;
}
int orc_preconditioner_on(void*) { return 0; }

int orc_erf_kernels(double a, double r, int nmax, double* B) {
    return guarded([&] { pgmcore::erfKernels(a, r, nmax, B); });
}

int orc_erfc_kernels(double a, double r, int nmax, double* B) {
    return guarded([&] { pgmcore::erfcKernels(a, r, nmax, B); });
}

int orc_pair_kernels(double a, double beta, double r, int nmax, double* G) {
    return guarded([&] { pgmcore::ewaldPairKernels(a, beta, r, nmax, G); });
}

int orc_gvdw_g(double y, double* G, double* H) {
    return guarded([&] { pgmcore::gvdwG(y, *G, *H); });
}

int orc_gvdw_pair(double r, double bij, double A, double C6, double B, int rep, double* energy, double* dEdR) {
    return guarded([&] { *energy = pgmcore::gvdwPair(r, bij, A, C6, B, rep, *dEdR); });
}

int orc_gvdw_sum(int, const double*, const double*, const double*, const double*, const double*, const double*, int,
                 double, int, const int*, const int*, double*, double*, double*) {
    return guarded([&] {
        throw std::runtime_error("pgm_ref: the reference core has no GVDW sum (core/gvdw_kernel.h has the pair only)");
    });
}

int orc_bspline_moduli(int K, int p, double* out) {
    return guarded([&] {
        const std::vector<double> m = pgmcore::bsplineModuli(K, p);
        std::memcpy(out, m.data(), K*sizeof(double));
    });
}

double orc_ewald_coefficient(double rc, double rtol) { return pgmcore::ewaldCoefficient(rc, rtol); }

int orc_fft3d_r2c(int K1, int K2, int K3, const double* in, double* out) {
    return guarded([&] {
        const size_t K3h = K3/2+1;
        pocketfft::shape_t shape = {(size_t) K1, (size_t) K2, (size_t) K3};
        pocketfft::stride_t sr = {(ptrdiff_t) (sizeof(double)*K2*K3), (ptrdiff_t) (sizeof(double)*K3),
                                  (ptrdiff_t) sizeof(double)};
        pocketfft::stride_t sc = {(ptrdiff_t) (2*sizeof(double)*K2*K3h), (ptrdiff_t) (2*sizeof(double)*K3h),
                                  (ptrdiff_t) (2*sizeof(double))};
        pocketfft::r2c(shape, sr, sc, {0, 1, 2}, pocketfft::FORWARD, in, reinterpret_cast<std::complex<double>*>(out),
                       1.0);
    });
}

int orc_fft3d_c2r(int K1, int K2, int K3, const double* in, double* out) {
    return guarded([&] {
        const size_t K3h = K3/2+1;
        pocketfft::shape_t shape = {(size_t) K1, (size_t) K2, (size_t) K3};
        pocketfft::stride_t sr = {(ptrdiff_t) (sizeof(double)*K2*K3), (ptrdiff_t) (sizeof(double)*K3),
                                  (ptrdiff_t) sizeof(double)};
        pocketfft::stride_t sc = {(ptrdiff_t) (2*sizeof(double)*K2*K3h), (ptrdiff_t) (2*sizeof(double)*K3h),
                                  (ptrdiff_t) (2*sizeof(double))};
        pocketfft::c2r(shape, sc, sr, {0, 1, 2}, pocketfft::BACKWARD,
                       reinterpret_cast<const std::complex<double>*>(in), out, 1.0);
    });
}

}   // extern "C"
