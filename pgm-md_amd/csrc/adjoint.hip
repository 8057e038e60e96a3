// Per-atom stages of the parameter gradients (FP64, sorted order inside, caller's order out).
//
// The energy E(r, mu; theta) is stationary in the induced dipoles mu, so for the forces F = -dE/dr|mu and a loss
// L(E, F) with cotangents wE = dL/dE, g = dL/dF:
//   dL/dtheta = wE dE/dtheta - D_(g, -lambda) (dE/dtheta)
// where D_(g, -lambda) is the directional derivative when every atom moves by g (the covalent dipoles follow
// the geometry) and mu changes by -lambda, and lambda solves the induction system A lambda = -E'(g) with the
// symmetric matrix A = alpha^-1 - T of the induced-dipole solve and E'(g) the change of the total field along g
// at fixed mu.  The parameter derivatives at fixed dipoles (reference energy terms, core/pgm_core.cpp:422-534):
//   dE/dq_i     = ke (phi_i - 2 beta/sqrt(pi) q_i - pi Q/(V beta^2))     phi: direct + reciprocal potential
//   dE/dc_m     = -ke E_a . u_m                                          u_m: unit vector atom -> partner
//   dE/dalpha_i = -ke |mu_i|^2 / (2 alpha_i^2)
//   dE/dR_i     = ke sum_j dU_ij/da (-2 R_i a^3)                          (pair pass)
#include "kernels.hpp"

namespace pgmhip {

namespace {

inline int nblk(long long threads, int bs) { return (int) ((threads+bs-1)/bs); }

// Change of the permanent dipoles of the sorted atoms when the atoms move by disp (caller's order):
// p_i = sum_m c_m v_m/|v_m|, v_m = r_partner - r_i, so p'_i = sum_m c_m (v'_m - (u_m . v'_m) u_m)/|v_m|.
template<bool BATCH>
__global__ void kPermTangent(int n, const int* order, const double* pos, const double* disp, const int* covStart,
                             const int* covPartner, const double* covStrength, BoxSel<double> bs, double* pdot) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const BoxT<double> box = selectBox<BATCH>(bs, s);
    const int i = order[s];
    double p[3] = {0, 0, 0};
    for (int m = covStart[i]; m < covStart[i+1]; m++) {
        const int j = covPartner[m];
        double x = pos[3*j]-pos[3*i], y = pos[3*j+1]-pos[3*i+1], z = pos[3*j+2]-pos[3*i+2];
        minImage(x, y, z, box);
        const double len = sqrt(x*x+y*y+z*z);
        const double u[3] = {x/len, y/len, z/len};
        const double v[3] = {disp[3*j]-disp[3*i], disp[3*j+1]-disp[3*i+1], disp[3*j+2]-disp[3*i+2]};
        const double uv = u[0]*v[0]+u[1]*v[1]+u[2]*v[2];
        const double c = covStrength[m]/len;
#pragma unroll
        for (int k = 0; k < 3; k++)
            p[k] += c*(v[k]-uv*u[k]);
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        pdot[3*s+k] = p[k];
}

__global__ void kSort3(int n, const int* order, const double* v, double* out) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= 3*n)
        return;
    const int s = k/3, c = k-3*s;
    out[k] = v[3*order[s]+c];
}

// Right-hand side of the adjoint solve: -(E'_pair + E'_rec + cself p') on polarizable rows.
__global__ void kAdjointRhs(int n, const double* alpha, const double* pairOut, const double* pmeOut,
                            const double* pdot, double cself, double* b) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= 3*n)
        return;
    const int s = k/3;
    b[k] = alpha[s] > 0 ? -(pairOut[6*n+k]+pmeOut[6*n+k]+cself*pdot[k]) : 0.0;
}

// d' = p' - lambda (lambda may be null: no induced dipoles), as FP64 and as the pair/PME vector type.
template<class T>
__global__ void kDipoleTangent(int n, const double* pdot, const double* lambda, double* ddot,
                               typename Vec4<T>::type* ddotT) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        v[c] = pdot[3*s+c]-(lambda ? lambda[3*s+c] : 0.0);
        ddot[3*s+c] = v[c];
    }
    typename Vec4<T>::type o;
    o.x = (T) v[0];
    o.y = (T) v[1];
    o.z = (T) v[2];
    o.w = T(0);
    ddotT[s] = o;
}

__global__ void kParamAtoms(ParamAtomArgs a) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= a.n)
        return;
    const int n = a.n;
    const int i = a.order[s];
    const double ke = a.ke;
    const bool tan = a.ddot != nullptr;
    // charges
    const double phi = a.pairOut[s]+a.pmeOut[s];
    const double qBack = a.sysOf ? a.qBackSys[a.sysOf[s]] : a.qBackground;
    const double dq = ke*(phi-a.qSelf*a.q[s]-qBack);
    const double wE = a.sysOf && a.wESys ? a.wESys[a.sysOf[s]] : a.wE;
    double gq = wE*dq;
    if (tan)
        gq -= ke*(a.pairOut[5*n+s]+a.pmeOut[5*n+s]);
    // polarizabilities
    const double al = a.alpha[s];
    double galpha = 0;
    if (al > 0) {
        double mm = 0, ml = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            mm += a.mu[3*s+c]*a.mu[3*s+c];
            if (a.lambda)
                ml += a.mu[3*s+c]*a.lambda[3*s+c];
        }
        galpha = wE*(-ke*mm/(2*al*al));
        if (tan)
            galpha -= ke*ml/(al*al);
    }
    // radii
    double gR = wE*ke*a.pairOut[4*n+s];
    if (tan)
        gR -= ke*a.pairOut[9*n+s];
    if (a.gq)
        a.gq[i] = gq;
    if (a.galpha)
        a.galpha[i] = galpha;
    if (a.gR)
        a.gR[i] = gR;
    // total field and its tangent at the atom, for the covalent strengths
#pragma unroll
    for (int c = 0; c < 3; c++) {
        a.etot[3*s+c] = a.pairOut[n+3*s+c]+a.pmeOut[n+3*s+c]+a.cself*a.d[3*s+c];
        if (tan)
            a.etotDot[3*s+c] = a.pairOut[6*n+3*s+c]+a.pmeOut[6*n+3*s+c]+a.cself*a.ddot[3*s+c];
    }
}

// One covalent entry (CSR order k, written at its input index): dL/dc = -wE ke E.u + ke (E'.u + E.u').
template<bool BATCH>
__global__ void kParamCov(int ncov, const int* covAtomOf, const int* covPartner, const int* inputOf, const int* rank,
                          const double* pos, const double* disp, const double* etot, const double* etotDot,
                          BoxSel<double> bs, double ke, double wE0, const double* wESys, double* gc) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= ncov)
        return;
    const int i = covAtomOf[k], j = covPartner[k];
    const BoxT<double> box = selectBox<BATCH>(bs, i);
    const double wE = BATCH && wESys ? wESys[bs.sysOf[i]] : wE0;
    double x = pos[3*j]-pos[3*i], y = pos[3*j+1]-pos[3*i+1], z = pos[3*j+2]-pos[3*i+2];
    minImage(x, y, z, box);
    const double len = sqrt(x*x+y*y+z*z);
    const double u[3] = {x/len, y/len, z/len};
    const int s = rank[i];
    const double* E = etot+3*s;
    double g = -wE*ke*(E[0]*u[0]+E[1]*u[1]+E[2]*u[2]);
    if (disp) {
        const double v[3] = {disp[3*j]-disp[3*i], disp[3*j+1]-disp[3*i+1], disp[3*j+2]-disp[3*i+2]};
        const double uv = u[0]*v[0]+u[1]*v[1]+u[2]*v[2];
        const double* Ed = etotDot+3*s;
        double eu = 0;
#pragma unroll
        for (int c = 0; c < 3; c++)
            eu += Ed[c]*u[c]+E[c]*(v[c]-uv*u[c])/len;
        g += ke*eu;
    }
    gc[inputOf[k]] = g;
}

}   // namespace

void launchPermTangent(hipStream_t st, int n, const int* order, const double* pos, const double* disp,
                       const int* covStart, const int* covPartner, const double* covStrength, BoxSel<double> box,
                       double* pdot) {
    if (box.table)
        kPermTangent<true><<<nblk(n, 256), 256, 0, st>>>(n, order, pos, disp, covStart, covPartner, covStrength, box,
                                                         pdot);
    else
        kPermTangent<false><<<nblk(n, 256), 256, 0, st>>>(n, order, pos, disp, covStart, covPartner, covStrength, box,
                                                          pdot);
}

void launchSort3(hipStream_t st, int n, const int* order, const double* v, double* out) {
    kSort3<<<nblk(3LL*n, 256), 256, 0, st>>>(n, order, v, out);
}

void launchAdjointRhs(hipStream_t st, int n, const double* alpha, const double* pairOut, const double* pmeOut,
                      const double* pdot, double cself, double* b) {
    kAdjointRhs<<<nblk(3LL*n, 256), 256, 0, st>>>(n, alpha, pairOut, pmeOut, pdot, cself, b);
}

template<class T>
void launchDipoleTangent(hipStream_t st, int n, const double* pdot, const double* lambda, double* ddot,
                         typename Vec4<T>::type* ddotT) {
    kDipoleTangent<T><<<nblk(n, 256), 256, 0, st>>>(n, pdot, lambda, ddot, ddotT);
}

void launchParamAtoms(hipStream_t st, const ParamAtomArgs& a) {
    kParamAtoms<<<nblk(a.n, 256), 256, 0, st>>>(a);
}

void launchParamCov(hipStream_t st, int ncov, const int* covAtomOf, const int* covPartner, const int* inputOf,
                    const int* rank, const double* pos, const double* disp, const double* etot,
                    const double* etotDot, BoxSel<double> box, double ke, double wE, const double* wESys,
                    double* gc) {
    if (ncov <= 0)
        return;
    if (box.table)
        kParamCov<true><<<nblk(ncov, 256), 256, 0, st>>>(ncov, covAtomOf, covPartner, inputOf, rank, pos, disp, etot,
                                                         etotDot, box, ke, wE, wESys, gc);
    else
        kParamCov<false><<<nblk(ncov, 256), 256, 0, st>>>(ncov, covAtomOf, covPartner, inputOf, rank, pos, disp,
                                                          etot, etotDot, box, ke, wE, wESys, gc);
}

template void launchDipoleTangent<float>(hipStream_t, int, const double*, const double*, double*, float4*);
template void launchDipoleTangent<double>(hipStream_t, int, const double*, const double*, double*, double4*);

}   // namespace pgmhip
