// The reference's C++ APIs on the GPU through include/pgm_hip.hpp, against the FP64 oracle (oracle/liborc.so):
// random water-like molecules in the triclinic box of core/tests/test_core.cpp, pgmcore-style (key rule,
// strain) and pgmcuda-style (computeF + forces()) calls.  Prints one line per check; exit status 0 = pass.
// Built by __graft_entry__.build() into tests/cpp/test_facade (g++, linked to libpgm_hip.so and liborc.so).
#include "pgm_hip.hpp"
#include "pgm_oracle.h"

#include <cmath>
#include <cstdio>
#include <random>

using pgmhip::core::V3;

int main() {
    const V3 box[3] = {{2.0, 0, 0}, {-0.5, 1.9, 0}, {-0.4, -0.7, 1.8}};
    std::mt19937 rng(11);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<V3> pos;
    pgmhip::core::Topology top;
    for (int m = 0; m < 24; m++) {
        const double f[3] = {U(rng), U(rng), U(rng)};
        V3 o;
        for (int k = 0; k < 3; k++)
            o[k] = f[0]*box[0][k]+f[1]*box[1][k]+f[2]*box[2][k];
        const int base = (int) pos.size();
        pos.push_back(o);
        for (int h = 0; h < 2; h++) {
            const double th = 2*M_PI*U(rng), ph = std::acos(2*U(rng)-1);
            pos.push_back(V3{o[0]+0.096*std::sin(ph)*std::cos(th), o[1]+0.096*std::sin(ph)*std::sin(th),
                             o[2]+0.096*std::cos(ph)});
        }
        top.q.insert(top.q.end(), {-0.8, 0.4, 0.4});
        top.radius.insert(top.radius.end(), {0.0605, 0.0536, 0.0536});
        top.alpha.insert(top.alpha.end(), {1.1e-3, 0.33e-3, 0.33e-3});
        for (int h = 1; h <= 2; h++) {
            top.covAtom.insert(top.covAtom.end(), {base, base+h});
            top.covPartner.insert(top.covPartner.end(), {base+h, base});
            top.covStrength.insert(top.covStrength.end(), {-0.019, 0.0086});
        }
    }
    const int n = (int) pos.size();
    pgmhip::core::Settings s;
    s.cutoff = 0.85;
    s.beta = 5.0;
    s.grid[0] = s.grid[1] = s.grid[2] = 48;
    s.tol = 1e-12;
    s.maxIter = 1000;
    s.peek = 0;
    s.predictor = 0;
    pgmhip::core::Engine gpu(top, s);
    std::vector<V3> F(n, V3{0, 0, 0});
    const double e = gpu.compute(pos, box, &F);

    orc_settings os{1, s.cutoff, s.beta, {48, 48, 48}, 6, 1, s.tol, s.maxIter, 0.0, 0, s.ke};
    void* o = orc_create(&os, n, top.q.data(), top.radius.data(), top.alpha.data(), (int) top.covAtom.size(),
                         top.covAtom.data(), top.covPartner.data(), top.covStrength.data());
    std::vector<double> fo(3*n, 0.0);
    double eo = 0;
    double b9[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            b9[3*r+c] = box[r][c];
    orc_compute(o, &pos[0][0], b9, fo.data(), &eo);
    double so[9];
    orc_strain(o, so);
    double df = 0, fmax = 0, ds = 0, smax = 0;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            df = std::max(df, std::fabs(F[i][k]-fo[3*i+k]));
            fmax = std::max(fmax, std::fabs(fo[3*i+k]));
        }
    for (int k = 0; k < 9; k++) {
        ds = std::max(ds, std::fabs(gpu.strainDerivative()[k]-so[k]));
        smax = std::max(smax, std::fabs(so[k]));
    }
    int fails = 0;
    const bool okE = std::fabs(e-eo) <= 1e-11*std::fabs(eo), okF = df <= 1e-9*fmax, okS = ds <= 1e-9*smax;
    printf("core::Engine FP64: E %.10f oracle %.10f (%s); max|dF| %.2e (%s); max|dS| %.2e (%s); %d iterations\n", e,
           eo, okE ? "ok" : "FAIL", df, okF ? "ok" : "FAIL", ds, okS ? "ok" : "FAIL", gpu.iterations());
    fails += !okE+!okF+!okS;
    // pgmcore::Engine::field of other charges and dipoles at the geometry of the last compute
    {
        std::vector<double> q2(n);
        std::vector<pgmhip::core::V3> d2(n);
        for (int i = 0; i < n; i++) {
            q2[i] = 0.3*std::sin(1.7*i);
            for (int k = 0; k < 3; k++)
                d2[i][k] = 0.01*std::cos(0.9*i+2.1*k);
        }
        const std::vector<pgmhip::core::V3> Eg = gpu.field(q2, d2);
        std::vector<double> Eo(3*(size_t) n);
        orc_field(o, q2.data(), &d2[0][0], Eo.data());
        double dE = 0, Emax = 0;
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                dE = std::max(dE, std::fabs(Eg[i][k]-Eo[3*i+k]));
                Emax = std::max(Emax, std::fabs(Eo[3*i+k]));
            }
        const bool okField = dE <= 1e-11*Emax;
        printf("core::Engine::field: max|dE| %.2e of %.2e (%s)\n", dE, Emax, okField ? "ok" : "FAIL");
        fails += !okField;
    }
    // same key: starts from the solution, no iterations
    gpu.setHistoryKey(7);
    gpu.compute(pos, box, nullptr);
    gpu.compute(pos, box, nullptr);
    const bool okKey = gpu.iterations() <= 1;   // starts from the converged solution of the same key
    printf("core::Engine same history key: %d iterations (%s)\n", gpu.iterations(), okKey ? "ok" : "FAIL");
    fails += !okKey;
    // pgmcuda-style engine, mixed precision, float positions
    pgmhip::Settings g;
    g.cutoff = 0.85;
    g.beta = 5.0;
    g.grid[0] = g.grid[1] = g.grid[2] = 48;
    g.tol = 1e-8;
    g.skin = 0.0;
    pgmhip::Engine mix(g);
    mix.setAtoms(n, top.q.data(), top.radius.data(), top.alpha.data(), nullptr, n);
    mix.setCovalent((int) top.covAtom.size(), top.covAtom.data(), top.covPartner.data(), top.covStrength.data());
    std::vector<float> pf(3*n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++)
            pf[3*i+k] = (float) pos[i][k];
    double bb[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            bb[r][c] = box[r][c];
    const pgmhip::Terms t = mix.computeF(pf.data(), bb, true, nullptr);
    double num = 0, den = 0;
    for (int k = 0; k < 3*n; k++) {
        num += (mix.forces()[k]-fo[k])*(mix.forces()[k]-fo[k]);
        den += fo[k]*fo[k];
    }
    const double rel = std::sqrt(num/den);
    const bool okM = rel < 5e-5 && std::fabs(t.total()/eo-1) < 1e-6;
    printf("Engine mixed computeF: force rel-RMS %.2e, energy rel %.2e (%s)\n", rel, t.total()/eo-1, okM ? "ok" : "FAIL");
    fails += !okM;
    // pgmcuda-style engine in FP64 over the caller's pair list (setCandidates): every pair within rc + 0.05 nm at
    // its shortest image, found among the 125 images t in [-2, 2]^3, one segment per pair
    std::vector<int> segI, segShift, segBegin, candJ;
    for (int i = 0; i < n; i++)
        for (int j = i+1; j < n; j++) {
            double best = 1e30;
            int bt[3] = {0, 0, 0};
            for (int t0 = -2; t0 <= 2; t0++)
                for (int t1 = -2; t1 <= 2; t1++)
                    for (int t2 = -2; t2 <= 2; t2++) {
                        double r2 = 0;
                        for (int k = 0; k < 3; k++) {
                            const double x = pos[i][k]-pos[j][k]+t0*box[0][k]+t1*box[1][k]+t2*box[2][k];
                            r2 += x*x;
                        }
                        if (r2 < best) {
                            best = r2;
                            bt[0] = t0;
                            bt[1] = t1;
                            bt[2] = t2;
                        }
                    }
            if (best < (s.cutoff+0.05)*(s.cutoff+0.05)) {
                segBegin.push_back((int) candJ.size());
                segI.push_back(i);
                segShift.insert(segShift.end(), {bt[0], bt[1], bt[2]});
                candJ.push_back(j);
            }
        }
    segBegin.push_back((int) candJ.size());
    pgmhip::Settings d = g;
    d.mixed = false;
    d.tol = 1e-12;
    d.maxIter = 1000;
    d.peek = 0;
    d.predictor = 0;
    pgmhip::Engine ext(d);
    ext.setAtoms(n, top.q.data(), top.radius.data(), top.alpha.data(), nullptr, n);
    ext.setCovalent((int) top.covAtom.size(), top.covAtom.data(), top.covPartner.data(), top.covStrength.data());
    ext.setCandidates((int) segI.size(), segI.data(), segShift.data(), segBegin.data(), candJ.data());
    std::vector<double> fe(3*n, 0.0);
    const pgmhip::Terms te = ext.compute(&pos[0][0], bb, true, fe.data(), nullptr);
    double dfe = 0;
    for (int k = 0; k < 3*n; k++)
        dfe = std::max(dfe, std::fabs(fe[k]-fo[k]));
    const bool okX = std::fabs(te.total()-eo) <= 1e-11*std::fabs(eo) && dfe <= 1e-9*fmax;
    printf("Engine FP64 over setCandidates (%zu pairs): E %.10f oracle %.10f, max|dF| %.2e (%s)\n", candJ.size(),
           te.total(), eo, dfe, okX ? "ok" : "FAIL");
    fails += !okX;
    orc_destroy(o);
    printf(fails ? "FAILED\n" : "PASSED\n");
    return fails;
}
