// Device helpers shared by the pGM HIP kernels (gfx950): the box and its minimum image, the Gaussian-damped
// Ewald pair kernels G_n, cardinal B-splines, sub-wavefront reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pgmhip {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPi = 1.7724538509055160273;
constexpr double kTwoOverSqrtPi = 1.1283791670955125739;

// Reduced triclinic box (rows a = (ax,0,0), b = (bx,by,0), c = (cx,cy,cz)) and its inverse
// recip[k][a] = (H^-1)_{k a}: the fractional coordinate a of r is sum_k r_k recip[k][a].
template<class T>
struct BoxT {
    T ax, bx, by, cx, cy, cz;
    T iax, iby, icz;   // 1/ax, 1/by, 1/cz
    T recip[3][3];
};

// Minimum image in a reduced triclinic box: c first, then b, then a (core/pgm_math.h:153-160).  rint of a
// product instead of round of a quotient: they differ only at half-box ties, far outside any pair cutoff.
template<class T>
__device__ __forceinline__ void minImage(T& x, T& y, T& z, const BoxT<T>& b) {
    T k = rint(z*b.icz);
    x -= k*b.cx;
    y -= k*b.cy;
    z -= k*b.cz;
    k = rint(y*b.iby);
    x -= k*b.bx;
    y -= k*b.by;
    k = rint(x*b.iax);
    x -= k*b.ax;
}

// The gfx950 dispatcher deals workgroups to its 8 XCDs round-robin (workgroup b runs on XCD b % 8), and each
// XCD has an L2 of its own.  This renumbering gives XCD x a contiguous range of logical blocks, so the kernels
// over spatially sorted atoms gather their partners from one slab of the system per L2 instead of from all of
// it.  A bijection of [0, gridDim.x) for any grid size.  reverse: each XCD walks its range from the end (the
// part of the data that the pass before, walking forward, touched last comes first).
__device__ __forceinline__ int xcdBlock(bool reverse = false) {
    constexpr int kXcd = 8;
    const int nb = (int) gridDim.x, b = (int) blockIdx.x;
    const int per = nb/kXcd, rem = nb%kXcd, x = b%kXcd, local = b/kXcd;
    const int l = reverse ? (x < rem ? per : per-1)-local : local;
    return x < rem ? x*(per+1)+l : rem*(per+1)+(x-rem)*per+l;
}

// Precision traits of the pair kernels: series length below a r = 1 and the damping cut beyond which the
// erfc(a r) part is dropped (the reference's GPU constants, openmm/platforms/common/src/CommonPGMKernels.cpp:351-352).
template<class T>
struct PairPrec;
template<>
struct PairPrec<float> {
    static constexpr int kSeries = 12;
    static constexpr float kDampCut = 4.5f;
};
template<>
struct PairPrec<double> {
    static constexpr int kSeries = 15;
    static constexpr double kDampCut = 6.0;
};

__device__ __forceinline__ float erfcT(float x) { return erfcf(x); }
__device__ __forceinline__ double erfcT(double x) { return erfc(x); }
__device__ __forceinline__ float erfT(float x) { return erff(x); }
__device__ __forceinline__ double erfT(double x) { return erf(x); }
__device__ __forceinline__ float expT(float x) { return __expf(x); }
__device__ __forceinline__ double expT(double x) { return exp(x); }
__device__ __forceinline__ float rsqrtT(float x) { return __frsqrt_rn(x); }
__device__ __forceinline__ double rsqrtT(double x) { return rsqrt(x); }

// exp(x^2) erfc(x) for 0 <= x <= 4.5 in FP32 (the range of the pair kernels): t Q(t) with t = 1/(1 + x/2) and
// Q a degree-8 polynomial fitted to exp(x^2) erfc(x)/t on that range (minimax relative error by iteratively
// reweighted least squares, pgm-md_amd/csrc/erfcx_fit.py; 1.6e-8 for the fit, below 3e-7 evaluated in FP32).
// erfc(x) is this times exp(-x^2), the factor the kernels' Gaussian term needs too, so that a kernel evaluation
// takes one exp per width instead of two.  v_rcp instead of a division.
__device__ __forceinline__ float erfcxFast(float x) {
    const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, x, 1.0f));
    float p = -4.658861831e-02f;
    p = fmaf(t, p, 2.203130424e-01f);
    p = fmaf(t, p, -3.455681503e-01f);
    p = fmaf(t, p, 1.038070619e-01f);
    p = fmaf(t, p, 1.041026637e-01f);
    p = fmaf(t, p, 1.415342391e-01f);
    p = fmaf(t, p, 2.605046332e-01f);
    p = fmaf(t, p, 2.796235681e-01f);
    p = fmaf(t, p, 2.822715342e-01f);
    return t*p;
}

// erfc(x) for 0 <= x <= 4.5 in FP32: erfcxFast(x) exp(-x^2); relative error below 2.5e-6 in FP32 (the rounding
// of x^2 in the exponent dominates near x = 4.5).
__device__ __forceinline__ float erfcFast(float x) { return erfcxFast(x)*__expf(-x*x); }

// B_n of erf(a r)/r, n < NMAX, by the series in s = (a r)^2 (valid for s < 1; no cancellation at short
// range): B_n = (2a/sqrt(pi)) (2a^2)^n sum_m (-s)^m / (m! (2m+2n+1)) (core/pgm_math.h:31-49, docs/spec.md Sec. 4).
template<class T, int NMAX>
__device__ __forceinline__ void erfSeries(T a, T s, T* B) {
    constexpr int NS = PairPrec<T>::kSeries;
    T term[NS];
    term[0] = T(1);
#pragma unroll
    for (int m = 1; m < NS; m++)
        term[m] = -term[m-1]*s*(T(1)/T(m));
    T scale = T(kTwoOverSqrtPi)*a;
    const T twoA2 = T(2)*a*a;
#pragma unroll
    for (int k = 0; k < NMAX; k++) {
        T acc = T(0);
#pragma unroll
        for (int m = NS-1; m >= 0; m--)
            acc += term[m]*(T(1)/T(2*m+2*k+1));
        B[k] = scale*acc;
        scale *= twoA2;
    }
}

// B_n of erf(a r)/r by the closed-form upward recursion (s >= 1), or by the series below.
template<class T, int NMAX>
__device__ __forceinline__ void erfKernels(T a, T r, T ir2, T* B) {
    const T x = a*r, s = x*x;
    if (s < T(1)) {
        erfSeries<T, NMAX>(a, s, B);
        return;
    }
    const T g = T(kTwoOverSqrtPi)*a*expT(-s);
    B[0] = erfT(x)/r;
    T scale = T(1);
    const T twoA2 = T(2)*a*a;
#pragma unroll
    for (int k = 1; k < NMAX; k++) {
        B[k] = (T(2*k-1)*B[k-1]-scale*g)*ir2;
        scale *= twoA2;
    }
}

// B_n of erfc(a r)/r: all terms positive (core/pgm_math.h:63-71).
template<class T, int NMAX>
__device__ __forceinline__ void erfcKernels(T a, T r, T ir2, T* B) {
    const T x = a*r;
    const T g = T(kTwoOverSqrtPi)*a*expT(-x*x);
    B[0] = erfcT(x)/r;
    T scale = T(1);
    const T twoA2 = T(2)*a*a;
#pragma unroll
    for (int k = 1; k < NMAX; k++) {
        B[k] = (T(2*k-1)*B[k-1]+scale*g)*ir2;
        scale *= twoA2;
    }
}

// FP32 forms of erfKernels / erfcKernels: 1/r from v_rsq, erfc(x) from erfcxFast(x) and the exp(-x^2) of the
// Gaussian term, no divisions.
template<int NMAX>
__device__ __forceinline__ void erfKernelsF(float a, float r, float rinv, float* B) {
    const float x = a*r, s = x*x;
    if (s < 1.0f) {
        erfSeries<float, NMAX>(a, s, B);
        return;
    }
    const float e = __expf(-s);
    const float g = float(kTwoOverSqrtPi)*a*e;
    const float ir2 = rinv*rinv;
    B[0] = (1.0f-erfcxFast(x)*e)*rinv;
    float scale = 1.0f;
    const float twoA2 = 2.0f*a*a;
#pragma unroll
    for (int k = 1; k < NMAX; k++) {
        B[k] = (float(2*k-1)*B[k-1]-scale*g)*ir2;
        scale *= twoA2;
    }
}

template<int NMAX>
__device__ __forceinline__ void erfcKernelsF(float a, float r, float rinv, float* B) {
    const float x = a*r;
    const float e = __expf(-x*x);   // shared by erfc(x) and the Gaussian term
    const float g = float(kTwoOverSqrtPi)*a*e;
    const float ir2 = rinv*rinv;
    B[0] = erfcxFast(x)*e*rinv;
    float scale = 1.0f;
    const float twoA2 = 2.0f*a*a;
#pragma unroll
    for (int k = 1; k < NMAX; k++) {
        B[k] = (float(2*k-1)*B[k-1]+scale*g)*ir2;
        scale *= twoA2;
    }
}

// G_n for FP32 from r^2 and rg_i + rg_j (a = (rg_i + rg_j)^-1/2): the branches of pairKernels below.
template<int NMAX>
__device__ __forceinline__ void pairKernelsF(float rgSum, float beta, float r2, float* G) {
    const float a = __builtin_amdgcn_rsqf(rgSum);
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2*rinv;
    float B[NMAX];
    if (a*r < 1.0f) {
        erfSeries<float, NMAX>(a, (a*r)*(a*r), G);
        erfKernelsF<NMAX>(beta, r, rinv, B);
    }
    else {
        erfcKernelsF<NMAX>(beta, r, rinv, G);
        if (a*r < PairPrec<float>::kDampCut)
            erfcKernelsF<NMAX>(a, r, rinv, B);
        else {
#pragma unroll
            for (int k = 0; k < NMAX; k++)
                B[k] = 0.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < NMAX; k++)
        G[k] -= B[k];
}

// G0 and G3 of a pair from its G1, G2 (FP32), by the recurrence of the erfc forms
// Bc_{n+1}[c] = ((2n+1) Bc_n[c] + (2c^2)^n h_c)/r^2, h_c = (2c/sqrt(pi)) exp(-c^2 r^2), applied to
// G_n = Bc_n[beta] - Bc_n[a]: G0 = r^2 G1 + h_a - h_beta, G3 = (5 G2 + (2 beta^2)^2 h_beta - (2a^2)^2 h_a)/r^2,
// with h_a = 0 beyond the damping cut as in pairKernelsF.  Close pairs (a r < 1, the series branch, where the
// upward recurrence would cancel) get all four kernels again from pairKernelsF.
__device__ __forceinline__ void g03FromG12(float G1, float G2, float r2, float rgSum, float beta, float& G0,
                                           float& G3);

__device__ __forceinline__ void g03FromG12(float G1, float G2, float r2, float rgSum, float beta, float& G0,
                                           float& G3) {
    const float a = __builtin_amdgcn_rsqf(rgSum);
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2*rinv;
    if (a*r < 1.0f) {
        float G[4];
        pairKernelsF<4>(rgSum, beta, r2, G);
        G0 = G[0];
        G3 = G[3];
        return;
    }
    const float ha = a*r < PairPrec<float>::kDampCut ? float(kTwoOverSqrtPi)*a*__expf(-a*a*r2) : 0.0f;
    const float hb = float(kTwoOverSqrtPi)*beta*__expf(-beta*beta*r2);
    const float ta = 2.0f*a*a, tb = 2.0f*beta*beta;
    G0 = r2*G1+ha-hb;
    G3 = (5.0f*G2+tb*tb*hb-ta*ta*ha)*(rinv*rinv);
}

// G_n = B_n[erf(a r)/r] - B_n[erf(beta r)/r] for n in [N0, NMAX): the series difference for a r < 1, the erfc
// form Bc(beta) - Bc(a) beyond (core/pgm_math.h:78-92), the Bc(a) part dropped beyond a r > kDampCut.
template<class T, int NMAX>
__device__ __forceinline__ void pairKernels(T a, T beta, T r, T* G) {
    const T ir2 = T(1)/(r*r);
    T B[NMAX];
    if (a*r < T(1)) {
        erfSeries<T, NMAX>(a, (a*r)*(a*r), G);
        erfKernels<T, NMAX>(beta, r, ir2, B);
    }
    else {
        erfcKernels<T, NMAX>(beta, r, ir2, G);
        if (a*r < PairPrec<T>::kDampCut)
            erfcKernels<T, NMAX>(a, r, ir2, B);
        else {
#pragma unroll
            for (int k = 0; k < NMAX; k++)
                B[k] = T(0);
        }
    }
#pragma unroll
    for (int k = 0; k < NMAX; k++)
        G[k] -= B[k];
}

// GVDW, the Gaussian-density van der Waals (core/gvdw_kernel.h): G(y) = F(y)/y^6 and H(y) = G'(y)/y with
// F = A^2 + g^2/2, A = erf y - (2y/sqrt(pi))(1 + 2y^2/3) e^{-y^2}, g = (4y^3/(3 sqrt(pi))) e^{-y^2}.  Below
// y = 0.6 the Taylor series of G in y^2 (coefficients c_k of G/C0, C0 = 8/(9 pi): exact rationals, from the
// series of erf and exp), above it the closed form.
template<class T>
__device__ __forceinline__ void gvdwG(T y, T& G, T& H) {
    constexpr double c[15] = {1.0, -2.0, 58.0/25, -188.0/105, 2222.0/2205, -1532.0/3465, 64244.0/405405,
                              -162056.0/3378375, 1587734.0/126351225, -211108.0/72747675,
                              11902564.0/19860115275.0, -611992.0/5465775315.0, 755028172.0/39529267903125.0,
                              -213335384.0/71152682225625.0, 15291334888.0/35078272337233125.0};
    const T c0 = T(8.0/(9.0*kPi)), s = y*y;
    if (y < T(0.6)) {
        T g = T(c[14]), h = T(28.0*c[14]);
#pragma unroll
        for (int k = 13; k >= 0; k--) {
            g = g*s+T(c[k]);
            if (k > 0)
                h = h*s+T(2.0*k*c[k]);
        }
        G = c0*g;
        H = c0*h;
        return;
    }
    const T e = expT(-s), isp = T(1.0/kSqrtPi);
    const T A = erfT(y)-T(2)*isp*y*(T(1)+T(2.0/3.0)*s)*e, g = T(4.0/3.0)*isp*s*y*e;
    const T dA = T(8.0/3.0)*isp*s*s*e, dg = T(4)*isp*s*(T(1)-T(2.0/3.0)*s)*e;
    const T F = A*A+T(0.5)*g*g, dF = T(2)*A*dA+g*dg, y6 = s*s*s;
    G = F/y6;
    H = (dF/y6-T(6)*F/(y6*y))/y;
}

// One GVDW pair: U = A exp(-B y^2) (Gaussian) or A exp(-B y) (Slater) - C6 b^6 G(y), y = b r; returns U and
// dEdR = -(1/r) dU/dr (the force on i is dEdR (r_i - r_j)); the Slater force at r = 0 is 0
// (core/gvdw_kernel.h:58-77).
template<class T>
__device__ __forceinline__ T gvdwPair(T r, T bij, T A, T C6, T B, bool slater, T& dEdR) {
    const T y = bij*r, b2 = bij*bij, b6 = b2*b2*b2;
    T erep, drep;
    if (!slater) {
        erep = A*expT(-B*y*y);
        drep = T(-2)*B*b2*erep;
    }
    else {
        erep = A*expT(-B*y);
        drep = r > T(0) ? -B*bij*erep/r : T(0);
    }
    T G, H;
    gvdwG<T>(y, G, H);
    dEdR = C6*b6*b2*H-drep;
    return erep-C6*b6*G;
}

// Cardinal B-spline weights of order P at fraction fr on the grid points (base - j), j = 0..P-1
// (core/pgm_core.cpp:131-152): th[j] = M_P(fr + j), dth[j] = M_P'(fr + j), d2th[j] = M_P''(fr + j).
template<class T, int P>
__device__ __forceinline__ void bsplines(T fr, T* th, T* dth, T* d2th) {
    T m[P+1];   // m[j] = M_k(fr + j), rebuilt for k = 1..P
    T m1[P+1], m2[P+1];
#pragma unroll
    for (int j = 0; j <= P; j++)
        m[j] = m1[j] = m2[j] = T(0);
    m[0] = T(1);
#pragma unroll
    for (int k = 2; k <= P; k++) {
        if (k == P-1) {
#pragma unroll
            for (int j = 0; j <= P; j++)
                m2[j] = m[j];   // M_{P-2}
        }
        if (k == P) {
#pragma unroll
            for (int j = 0; j <= P; j++)
                m1[j] = m[j];   // M_{P-1}
        }
        T nxt[P+1];
        const T inv = T(1)/T(k-1);
#pragma unroll
        for (int j = 0; j <= P; j++) {
            const T hi = (j < k-1 ? m[j] : T(0));
            const T lo = (j >= 1 && j-1 < k-1 ? m[j-1] : T(0));
            nxt[j] = (j < k ? ((fr+T(j))*hi+(T(k)-fr-T(j))*lo)*inv : T(0));
        }
#pragma unroll
        for (int j = 0; j <= P; j++)
            m[j] = nxt[j];
    }
#pragma unroll
    for (int j = 0; j < P; j++) {
        th[j] = m[j];
        const T a1 = (j < P-1 ? m1[j] : T(0)), b1 = (j >= 1 ? m1[j-1] : T(0));
        dth[j] = a1-b1;
        const T a2 = (j < P-2 ? m2[j] : T(0)), b2 = (j >= 1 && j-1 < P-2 ? m2[j-1] : T(0)),
                c2 = (j >= 2 && j-2 < P-2 ? m2[j-2] : T(0));
        d2th[j] = a2-T(2)*b2+c2;
    }
}

// Sum over the L lanes of a sub-wavefront group (L a power of two dividing 64).
template<int L, class V>
__device__ __forceinline__ V groupSum(V v) {
#pragma unroll
    for (int o = L/2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, L);
    return v;
}

// slot = atomicAdd(&count[key], 1) for every lane with valid set, with one atomic per distinct key in the
// wavefront: spatially sorted atoms share a key across most of a wavefront, and atomics on one address
// serialize.  Every lane of the wavefront calls it (valid = false where there is no item); the slots within a
// key follow lane order, and those of different wavefronts the order of their atomics.
__device__ __forceinline__ int waveSlot(int* count, int key, bool valid) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(valid);
    int slot = 0;
    while (pending != 0ull) {
        const int leader = __ffsll((unsigned long long) pending)-1;
        const int lk = __shfl(key, leader, 64);
        const unsigned long long same = __ballot(valid && key == lk) & pending;
        int base = 0;
        if (lane == leader)
            base = atomicAdd(&count[lk], __popcll(same));
        base = __shfl(base, leader, 64);
        if ((same >> lane) & 1ull)
            slot = base+__popcll(same & ((1ull << lane)-1ull));
        pending &= ~same;
    }
    return slot;
}

// Block-wide sum of a double (blockDim.x a multiple of 64, <= 1024); the result is valid in thread 0.
__device__ __forceinline__ double blockSum(double v, double* smem) {
    v = groupSum<64>(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        smem[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int) (blockDim.x >> 6); k++)
            r += smem[k];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double groupMax64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double blockMax(double v, double* smem) {
    v = groupMax64(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        smem[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int) (blockDim.x >> 6); k++)
            r = fmax(r, smem[k]);
    __syncthreads();
    return r;
}

// Vector types of the two precisions: xyz + charge (positions) or xyz + 0 (dipoles), 16 or 32 bytes.
template<class T>
struct Vec4;
template<>
struct Vec4<float> {
    typedef float4 type;
};
template<>
struct Vec4<double> {
    typedef double4 type;
};
template<class T>
struct Vec2;
template<>
struct Vec2<float> {
    typedef float2 type;
};
template<>
struct Vec2<double> {
    typedef double2 type;
};

}   // namespace pgmhip
