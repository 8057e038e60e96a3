// Smooth PME of charges and dipoles (Essmann 1995; dipoles as Sagui 2004) for pGM on gfx950.
//
//   bin/place  atoms binned into grid bricks once per evaluation, B-spline weights stored in brick order
//              (core/pgm_core.cpp:123-153 convention: M_p(fr + j) on grid point base - j);
//   spread     one workgroup per brick accumulates q theta theta theta + e . grad(theta theta theta) into an
//              FP64 LDS tile (brick + order-1 halo) with LDS atomics, then adds the tile to the grid
//              (core/pgm_core.cpp:191-207; replaces the reference's 64-bit fixed-point global atomics);
//   transform  in-house mixed-radix (4,2,3,5,7) Stockham FFTs in LDS: z (two real lines per complex
//              transform), y, then x fused with the influence function, the energy and the k-space strain
//              (core/pgm_core.cpp:213-237), and back;
//   gather     E = -grad phi at the atoms; at the end of an evaluation also the reciprocal forces
//              -[q grad phi + (grad grad phi) . d] (core/pgm_core.cpp:241-264, 453-487).
#include "kernels.hpp"

#include <type_traits>

#include <stdexcept>
#include <string>

namespace pgmhip {

FftPlan makePlan(int K) {
    FftPlan p{};
    p.K = K;
    int m = K;
    const int radices[5] = {4, 2, 3, 5, 7};
    for (int r : radices)
        while (m%r == 0) {
            if (p.nstage >= 10)
                throw std::runtime_error("pgm_hip: PME grid dimension has too many factors");
            p.radix[p.nstage++] = r;
            m /= r;
        }
    if (m != 1)
        throw std::runtime_error("pgm_hip: PME grid dimension " + std::to_string(K)+
                                 " must factor into 2, 3, 5 and 7");
    if (K >= 4096)
        throw std::runtime_error("pgm_hip: PME grid dimension " + std::to_string(K)+" above 4095");
    int ns = 1;
    for (int st = 0; st < p.nstage; st++) {
        p.byM[st] = fastDiv(K/p.radix[st]);
        p.byNs[st] = fastDiv(ns);
        ns *= p.radix[st];
    }
    return p;
}

namespace {

constexpr int kBlock = 256;

template<class T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
    T2 r;
    r.x = a.x*b.x-a.y*b.y;
    r.y = a.x*b.y+a.y*b.x;
    return r;
}

template<class T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) {
    T2 r;
    r.x = a.x+b.x;
    r.y = a.y+b.y;
    return r;
}

template<class T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) {
    T2 r;
    r.x = a.x-b.x;
    r.y = a.y-b.y;
    return r;
}

template<class T2>
__device__ __forceinline__ T2 conjIf(T2 a, bool c) {
    if (c)
        a.y = -a.y;
    return a;
}

// -i a (forward) or +i a (inverse)
template<class T2>
__device__ __forceinline__ T2 rotQuarter(T2 a, bool inverse) {
    T2 r;
    if (inverse) {
        r.x = -a.y;
        r.y = a.x;
    }
    else {
        r.x = a.y;
        r.y = -a.x;
    }
    return r;
}

// cos(2 pi k/R) and sin(2 pi k/R) for R = 5 or 7 and 0 < k < R (cos(2 pi/5) = (sqrt 5 - 1)/4, and so on).
__host__ __device__ constexpr double cosTwoPi(int R, int k) {
    const int q = k <= R/2 ? k : R-k;
    return R == 5 ? (q == 1 ? 0.30901699437494742410 : -0.80901699437494742410)
                  : (q == 1 ? 0.62348980185873353053 : q == 2 ? -0.22252093395631440429 : -0.90096886790241912624);
}
__host__ __device__ constexpr double sinTwoPi(int R, int k) {
    const int q = k <= R/2 ? k : R-k;
    const double s = R == 5 ? (q == 1 ? 0.95105651629515357212 : 0.58778525229247312917)
                            : (q == 1   ? 0.78183148246802980871
                               : q == 2 ? 0.97492791218182360702
                                        : 0.43388373911755812048);
    return k <= R/2 ? s : -s;
}

// One radix-R stage of a Stockham FFT over nlines lines of length K in LDS (line l at buf[l*K]): butterfly j of
// line l takes in[j + r*m], r < R (m = K/R), each times w^(r*kk*span) (kk = j mod Ns, span = K/(Ns*R)), and
// writes out[(j div Ns)*Ns*R + kk + s*Ns], s < R.  R is a compile-time constant, so the R inputs stay in
// registers: with a runtime radix the array was indexed at run time and lived in per-lane scratch memory.
template<int R, class T>
__device__ __forceinline__ void fftStage(const typename Vec2<T>::type* src, typename Vec2<T>::type* dst,
                                         const typename Vec2<T>::type* tw, int K, int Ns, FastDiv byM, FastDiv byNs,
                                         int nlines, bool inverse) {
    typedef typename Vec2<T>::type T2;
    const int m = K/R, span = K/(Ns*R);
    for (int idx = threadIdx.x; idx < nlines*m; idx += blockDim.x) {
        const int l = byM.div(idx), j = idx-l*m, jq = byNs.div(j), kk = j-jq*Ns;
        const T2* in = src+l*K;
        T2* out = dst+l*K+jq*Ns*R+kk;
        T2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const T2 a = in[j+r*m];
            v[r] = (r == 0 ? a : cmul(a, conjIf(tw[r*kk*span], inverse)));
        }
        if constexpr (R == 2) {
            out[0] = cadd(v[0], v[1]);
            out[Ns] = csub(v[0], v[1]);
        }
        else if constexpr (R == 4) {
            const T2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
            const T2 s13 = cadd(v[1], v[3]), d13 = rotQuarter(csub(v[1], v[3]), inverse);
            out[0] = cadd(s02, s13);
            out[Ns] = cadd(d02, d13);
            out[2*Ns] = csub(s02, s13);
            out[3*Ns] = csub(d02, d13);
        }
        else if constexpr (R == 3) {
            const T c = T(-0.5), sn = (inverse ? T(0.86602540378443864676) : T(-0.86602540378443864676));
            const T2 sum = cadd(v[1], v[2]), d = csub(v[1], v[2]);
            T2 t;
            t.x = v[0].x+c*sum.x;
            t.y = v[0].y+c*sum.y;
            out[0] = cadd(v[0], sum);
            T2 u;
            u.x = -sn*d.y;
            u.y = sn*d.x;
            out[Ns] = cadd(t, u);
            out[2*Ns] = csub(t, u);
        }
        else {
            // radix 5 and 7 in closed form: with a_r = v_r + v_{R-r}, b_r = v_r - v_{R-r} (r = 1..h, h = (R-1)/2),
            // X_s = v_0 + sum a_r cos(2 pi rs/R) -/+ i sum b_r sin(2 pi rs/R) and X_{R-s} the other sign
            // (forward/inverse); the cosines and sines fold to constants
            constexpr int h = (R-1)/2;
            T2 a[h+1], b[h+1];
            T2 x0 = v[0];
#pragma unroll
            for (int r = 1; r <= h; r++) {
                a[r] = cadd(v[r], v[R-r]);
                b[r] = csub(v[r], v[R-r]);
                x0 = cadd(x0, a[r]);
            }
            out[0] = x0;
#pragma unroll
            for (int s = 1; s <= h; s++) {
                T2 A = v[0], B;
                B.x = B.y = T(0);
#pragma unroll
                for (int r = 1; r <= h; r++) {
                    const T c = T(cosTwoPi(R, (r*s)%R)), sn = T(sinTwoPi(R, (r*s)%R));
                    A.x += c*a[r].x;
                    A.y += c*a[r].y;
                    B.x += sn*b[r].x;
                    B.y += sn*b[r].y;
                }
                const T2 rb = rotQuarter(B, inverse);
                out[s*Ns] = cadd(A, rb);
                out[(R-s)*Ns] = csub(A, rb);
            }
        }
    }
}

// Mixed-radix Stockham FFT of nlines lines of length K held in LDS (line l at buf[l*K]); tw[j] =
// e^{-2 pi i j/K} in LDS.  Returns the buffer that holds the result (natural order).
template<class T>
__device__ __forceinline__ typename Vec2<T>::type* ldsFft(typename Vec2<T>::type* src, typename Vec2<T>::type* dst,
                                                          const typename Vec2<T>::type* tw, const FftPlan& p,
                                                          int nlines, bool inverse) {
    typedef typename Vec2<T>::type T2;
    const int K = p.K;
    int Ns = 1;
    for (int st = 0; st < p.nstage; st++) {
        const int R = p.radix[st];
        const FastDiv byM = p.byM[st], byNs = p.byNs[st];
        switch (R) {   // makePlan's radices: 4, 2, 3, 5, 7
        case 2:
            fftStage<2, T>(src, dst, tw, K, Ns, byM, byNs, nlines, inverse);
            break;
        case 3:
            fftStage<3, T>(src, dst, tw, K, Ns, byM, byNs, nlines, inverse);
            break;
        case 4:
            fftStage<4, T>(src, dst, tw, K, Ns, byM, byNs, nlines, inverse);
            break;
        case 5:
            fftStage<5, T>(src, dst, tw, K, Ns, byM, byNs, nlines, inverse);
            break;
        default:
            fftStage<7, T>(src, dst, tw, K, Ns, byM, byNs, nlines, inverse);
            break;
        }
        __syncthreads();
        T2* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

// ---------------------------------------------------------------------------------------- binning

template<class T>
__device__ __forceinline__ void fracGrid(const double4& p, const BoxT<double>& box, const PmeGrid& g, int base[3],
                                         T fr[3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double t = p.x*box.recip[0][a]+p.y*box.recip[1][a]+p.z*box.recip[2][a];
        t -= floor(t);
        const double w = g.K[a]*t;
        int b = (int) floor(w);
        fr[a] = (T) (w-b);
        if (b >= g.K[a])
            b -= g.K[a];
        base[a] = b;
    }
}

// The grid and box of a system: the single one, or from the batch table.
template<bool BATCH>
__device__ __forceinline__ const PmeSys* sysEntry(const PmeSel& ps, int sys) {
    return BATCH ? ps.table+sys : nullptr;
}

template<class T, bool BATCH>
__global__ void kPmeBin(int n, const double4* posD, PmeSel ps, int* brickOf, int* slot, int* brickCount) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    const bool valid = s < n;   // no early return: the whole wavefront takes part in waveSlot
    int b = 0;
    if (valid) {
        const PmeSys* se = sysEntry<BATCH>(ps, BATCH ? ps.sysOf[s] : 0);
        const PmeGrid& g = BATCH ? se->g : ps.g;
        const BoxT<double>& box = BATCH ? se->box : ps.box;
        int base[3];
        T fr[3];
        fracGrid<T>(posD[s], box, g, base, fr);
        b = ((base[0]/g.brick)*g.nb[1]+base[1]/g.brick)*g.nb[2]+base[2]/g.brick+(BATCH ? se->brickOff : 0);
        brickOf[s] = b;
    }
    const int k = waveSlot(brickCount, b, valid);
    if (valid)
        slot[s] = k;
}

template<class T, int P, bool BATCH>
__global__ void kPmePlace(int n, const double4* posD, PmeSel ps, const int* brickOf, const int* slot,
                          const int* brickStart, int* pmeAtom, int* baseOut, T* spl) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const PmeSys* se = sysEntry<BATCH>(ps, BATCH ? ps.sysOf[s] : 0);
    const PmeGrid& g = BATCH ? se->g : ps.g;
    const BoxT<double>& box = BATCH ? se->box : ps.box;
    int base[3];
    T fr[3];
    fracGrid<T>(posD[s], box, g, base, fr);
    const int k = brickStart[brickOf[s]]+slot[s];
    pmeAtom[k] = s;
    T* o = spl+(size_t) k*9*P;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        baseOut[3*k+a] = base[a];
        T th[P], dth[P], d2th[P];
        bsplines<T, P>(fr[a], th, dth, d2th);
#pragma unroll
        for (int j = 0; j < P; j++) {
            o[a*3*P+j] = th[j];
            o[a*3*P+P+j] = dth[j];
            o[a*3*P+2*P+j] = d2th[j];
        }
    }
}

// ---------------------------------------------------------------------------------------- spreading

// TAN: instead of the charges and dipoles, the change of their grid density when every atom moves by g_s and its
// dipole changes by d'_s: the density q theta(r - r_s) + d . grad_s theta moves to first order by
// (q g + d') . grad_s theta + g^T (grad_s grad_s theta) d, i.e. a dipole q g + d' and a symmetric second moment
// (g d^T + d g^T)/2 on the second spline derivatives.
//
// The charge and dipole grids (not TAN) accumulate in 64-bit fixed point, GridFixed<T>::scale units per grid
// unit, in the LDS tile and in the global grid: integer additions commute, so the grid, and everything after
// it, is the same bit for bit whatever the order of the atomics (the reference's 2^32 fixed-point grid,
// core/cuda/pgm_cuda.cu:52,67-69,804-843).  Each contribution is rounded to the fixed-point unit once.  TAN
// grids, whose scale follows the caller's cotangent, accumulate in FP64 in the tile and in T in the grid.
template<class T, int P, bool TAN, bool BATCH>
__global__ void __launch_bounds__(kBlock) kSpread(PmeSel ps, const int* __restrict__ brickStart,
                                                  const int* __restrict__ pmeAtom, const int* __restrict__ baseArr,
                                                  const T* __restrict__ spl,
                                                  const typename Vec4<T>::type* __restrict__ dip,
                                                  const typename Vec4<T>::type* __restrict__ posq, int withQ,
                                                  T* __restrict__ gridAll, long long* __restrict__ fixedAll,
                                                  const int* skip,
                                                  const typename Vec4<T>::type* __restrict__ dipDot,
                                                  const typename Vec4<T>::type* __restrict__ disp) {
    typedef typename Vec4<T>::type T4;
    // The TAN tile accumulates in FP64 for both precisions: ds_add_f64 is several times faster than ds_add_f32
    // on gfx950, and the FP64 sums lose nothing.
    constexpr int W = GridFixed<T>::words;
    extern __shared__ unsigned char smemRaw[];
    double* tile = reinterpret_cast<double*>(smemRaw);
    unsigned long long* tileI = reinterpret_cast<unsigned long long*>(smemRaw);   // W words per point, apart
    if (skip && *skip)
        return;
    const int bid = blockIdx.x;
    const PmeSys* se = sysEntry<BATCH>(ps, BATCH ? ps.brickSys[bid] : 0);
    const PmeGrid& g = BATCH ? se->g : ps.g;
    const BoxT<double>& box = BATCH ? se->box : ps.box;
    T* __restrict__ grid = TAN ? gridAll+(BATCH ? se->gridOff : 0) : nullptr;
    long long* __restrict__ gridI = TAN ? nullptr : fixedAll+(BATCH ? se->gridOff : 0);
    const int lb = bid-(BATCH ? se->brickOff : 0);
    const int b2 = lb%g.nb[2], b1 = (lb/g.nb[2])%g.nb[1], b0 = lb/(g.nb[2]*g.nb[1]);
    const int E = g.brick+P-1, E3 = E*E*E;
    const int lo0 = b0*g.brick-(P-1), lo1 = b1*g.brick-(P-1), lo2 = b2*g.brick-(P-1);
    for (int i = threadIdx.x; i < (TAN ? 1 : W)*E3; i += blockDim.x)
        tileI[i] = 0ull;
    __syncthreads();
    T rc[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int a = 0; a < 3; a++)
            rc[k][a] = (T) (g.K[a]*box.recip[k][a]);
    const int first = brickStart[bid], count = brickStart[bid+1]-first;
    for (int it = threadIdx.x; it < count*P; it += blockDim.x) {
        const int k = first+it/P, j1 = it%P;
        const int s = pmeAtom[k];
        const T4 d = dip[s];
        const T q = (withQ || TAN) ? posq[s].w : T(0);
        T e[3], D[3], S[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
            e[a] = d.x*rc[0][a]+d.y*rc[1][a]+d.z*rc[2][a];
        if (TAN) {
            const T4 dd = dipDot[s], gs = disp[s];
            T ga[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                ga[a] = gs.x*rc[0][a]+gs.y*rc[1][a]+gs.z*rc[2][a];
                D[a] = q*ga[a]+dd.x*rc[0][a]+dd.y*rc[1][a]+dd.z*rc[2][a];
            }
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    S[a][b] = T(0.5)*(ga[a]*e[b]+ga[b]*e[a]);
        }
        const T* o = spl+(size_t) k*9*P;
        const T t1 = o[j1], u1 = o[P+j1], v1 = o[2*P+j1];
        const int l0 = baseArr[3*k]-j1-lo0, l1b = baseArr[3*k+1]-lo1, l2b = baseArr[3*k+2]-lo2;
        T t3[P], u3[P], v3[P];
#pragma unroll
        for (int j = 0; j < P; j++) {
            t3[j] = o[6*P+j];
            u3[j] = o[7*P+j];
            if (TAN)
                v3[j] = o[8*P+j];
        }
#pragma unroll
        for (int j2 = 0; j2 < P; j2++) {
            const T t2 = o[3*P+j2], u2 = o[4*P+j2];
            T A, C, V = 0;
            if (TAN) {
                const T v2 = o[5*P+j2];
                A = D[0]*u1*t2+D[1]*t1*u2+S[0][0]*v1*t2+S[1][1]*t1*v2+T(2)*S[0][1]*u1*u2;
                C = D[2]*t1*t2+T(2)*(S[0][2]*u1*t2+S[1][2]*t1*u2);
                V = S[2][2]*t1*t2;
            }
            else {
                A = (q*t1+e[0]*u1)*t2+e[1]*t1*u2;
                C = e[2]*t1*t2;
            }
            const int r0 = (l0*E+(l1b-j2))*E+l2b;
            if (TAN) {
#pragma unroll
                for (int j3 = 0; j3 < P; j3++)
                    atomicAdd(tile+r0-j3, (double) (A*t3[j3]+C*u3[j3]+V*v3[j3]));
            }
            else {
#pragma unroll
                for (int j3 = 0; j3 < P; j3++) {
                    long long hi, lo;
                    toFixed<T>(A*t3[j3]+C*u3[j3], hi, lo);
                    atomicAdd(tileI+r0-j3, (unsigned long long) hi);
                    if (W == 2)
                        atomicAdd(tileI+E3+r0-j3, (unsigned long long) lo);
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < E3; i += blockDim.x) {
        if (TAN ? tile[i] == 0.0 : (tileI[i] == 0ull && (W == 1 || tileI[E3+i] == 0ull)))
            continue;
        const int x2 = i%E, x1 = (i/E)%E, x0 = i/(E*E);
        int g0 = lo0+x0, g1 = lo1+x1, g2 = lo2+x2;
        g0 += (g0 < 0 ? g.K[0] : 0);
        g0 -= (g0 >= g.K[0] ? g.K[0] : 0);
        g1 += (g1 < 0 ? g.K[1] : 0);
        g1 -= (g1 >= g.K[1] ? g.K[1] : 0);
        g2 += (g2 < 0 ? g.K[2] : 0);
        g2 -= (g2 >= g.K[2] ? g.K[2] : 0);
        const size_t at = ((size_t) g0*g.K[1]+g1)*g.K[2]+g2;
        if (TAN)
            unsafeAtomicAdd(grid+at, (T) tile[i]);
        else {
            atomicAdd(reinterpret_cast<unsigned long long*>(gridI+at), tileI[i]);
            if (W == 2)
                atomicAdd(reinterpret_cast<unsigned long long*>(gridI+ps.fixedStride+at), tileI[E3+i]);
        }
    }
}

// ---------------------------------------------------------------------------------------- transforms

template<class T>
__device__ __forceinline__ void loadTwiddles(typename Vec2<T>::type* dst, const typename Vec2<T>::type* src, int K) {
    for (int i = threadIdx.x; i < K; i += blockDim.x)
        dst[i] = src[i];
}

// The group (grid shape) of this block in a merged launch of stage 0 (z), 1 (y) or 2 (x).
__device__ __forceinline__ int fftGroupOf(const FftGroup* groups, int ngroups, int stage) {
    int gi = 0;
    for (int k = 1; k < ngroups; k++)
        if ((int) blockIdx.x >= groups[k].blk0[stage])
            gi = k;
    return gi;
}

// z: two real lines per complex FFT, unpacked into their half spectra.
// Forward: from the fixed-point grid when fixedAll is set (each word zeroed once read: the grid is clean for
// the next spread), else from the real grid.
template<class T, bool INVERSE>
__global__ void __launch_bounds__(kBlock) kFftZ(const FftGroup* __restrict__ groups, int ngroups,
                                                T* __restrict__ realAll, typename Vec2<T>::type* __restrict__ cplxAll,
                                                const typename Vec2<T>::type* __restrict__ twAll, const int* skip,
                                                long long* __restrict__ fixedAll, size_t fixedStride) {
    typedef typename Vec2<T>::type T2;
    extern __shared__ unsigned char smemRaw[];
    if (skip && *skip)
        return;
    const FftGroup& grp = groups[fftGroupOf(groups, ngroups, 0)];
    const PmeGrid& g = grp.g;
    const int ngrid = grp.ngrid, pairsPerBlock = grp.pz, blk = (int) blockIdx.x-grp.blk0[0];
    T* __restrict__ real = realAll+grp.realOff;
    T2* __restrict__ cplx = cplxAll+grp.cplxOff;
    const FftPlan& p = g.plan[2];
    const int K = g.K[2], Kh = g.K3h, nlines = ngrid*g.K[0]*g.K[1];
    T2* tw = reinterpret_cast<T2*>(smemRaw);
    T2* b0 = tw+K;
    T2* b1 = b0+pairsPerBlock*K;
    loadTwiddles<T>(tw, twAll+grp.twOff+g.K[0]+g.K[1], K);
    const int pair0 = blk*pairsPerBlock;
    const int npairs = min(pairsPerBlock, (nlines+1)/2-pair0);
    if (!INVERSE) {
        long long* fixed = fixedAll ? fixedAll+grp.realOff : nullptr;
        for (int i = threadIdx.x; i < npairs*K; i += blockDim.x) {
            const int l = i/K, k = i-l*K;
            const int la = 2*(pair0+l), lb = la+1;
            T2 z;
            if (fixed) {
                z.x = takeFixed<T>(fixed+(size_t) la*K+k, fixedStride);
                z.y = (lb < nlines ? takeFixed<T>(fixed+(size_t) lb*K+k, fixedStride) : T(0));
            }
            else {
                z.x = real[(size_t) la*K+k];
                z.y = (lb < nlines ? real[(size_t) lb*K+k] : T(0));
            }
            b0[l*K+k] = z;
        }
    }
    else {
        for (int i = threadIdx.x; i < npairs*K; i += blockDim.x) {
            const int l = i/K, k = i-l*K;
            const int la = 2*(pair0+l), lb = la+1;
            const bool lower = k < Kh;
            const int kk = lower ? k : K-k;
            T2 X = cplx[(size_t) la*Kh+kk];
            T2 Y;
            Y.x = Y.y = T(0);
            if (lb < nlines)
                Y = cplx[(size_t) lb*Kh+kk];
            if (kk == 0 || (K%2 == 0 && kk == K/2)) {
                X.y = T(0);
                Y.y = T(0);
            }
            T2 z;
            if (lower) {
                z.x = X.x-Y.y;
                z.y = X.y+Y.x;
            }
            else {
                z.x = X.x+Y.y;
                z.y = Y.x-X.y;
            }
            b0[l*K+k] = z;
        }
    }
    __syncthreads();
    T2* res = ldsFft<T>(b0, b1, tw, p, npairs, INVERSE);
    if (!INVERSE) {
        for (int i = threadIdx.x; i < npairs*Kh; i += blockDim.x) {
            const int l = i/Kh, k = i-l*Kh;
            const int la = 2*(pair0+l), lb = la+1;
            const T2 Z = res[l*K+k], Zc = res[l*K+(K-k)%K];
            T2 X, Y;
            X.x = T(0.5)*(Z.x+Zc.x);
            X.y = T(0.5)*(Z.y-Zc.y);
            Y.x = T(0.5)*(Z.y+Zc.y);
            Y.y = T(-0.5)*(Z.x-Zc.x);
            cplx[(size_t) la*Kh+k] = X;
            if (lb < nlines)
                cplx[(size_t) lb*Kh+k] = Y;
        }
    }
    else {
        for (int i = threadIdx.x; i < npairs*K; i += blockDim.x) {
            const int l = i/K, k = i-l*K;
            const int la = 2*(pair0+l), lb = la+1;
            const T2 z = res[l*K+k];
            real[(size_t) la*K+k] = z.x;
            if (lb < nlines)
                real[(size_t) lb*K+k] = z.y;
        }
    }
}

// y: lines of K2 for fixed (x, z), a tile of consecutive z per block.
template<class T, bool INVERSE>
__global__ void __launch_bounds__(kBlock) kFftY(const FftGroup* __restrict__ groups, int ngroups,
                                                typename Vec2<T>::type* __restrict__ cplxAll,
                                                const typename Vec2<T>::type* __restrict__ twAll, const int* skip) {
    typedef typename Vec2<T>::type T2;
    extern __shared__ unsigned char smemRaw[];
    if (skip && *skip)
        return;
    const FftGroup& grp = groups[fftGroupOf(groups, ngroups, 1)];
    const PmeGrid& g = grp.g;
    const int zt = grp.zy, blk = (int) blockIdx.x-grp.blk0[1];
    T2* __restrict__ cplx = cplxAll+grp.cplxOff;
    const int K = g.K[1], Kh = g.K3h;
    const int ntile = (Kh+zt-1)/zt;
    const int x = blk/ntile, z0 = (blk%ntile)*zt;
    const int nl = min(zt, Kh-z0);
    T2* tw = reinterpret_cast<T2*>(smemRaw);
    T2* b0 = tw+K;
    T2* b1 = b0+zt*K;
    loadTwiddles<T>(tw, twAll+grp.twOff+g.K[0], K);
    for (int i = threadIdx.x; i < nl*K; i += blockDim.x) {
        const int y = i/nl, l = i-y*nl;
        b0[l*K+y] = cplx[((size_t) x*K+y)*Kh+z0+l];
    }
    __syncthreads();
    T2* res = ldsFft<T>(b0, b1, tw, g.plan[1], nl, INVERSE);
    for (int i = threadIdx.x; i < nl*K; i += blockDim.x) {
        const int y = i/nl, l = i-y*nl;
        cplx[((size_t) x*K+y)*Kh+z0+l] = res[l*K+y];
    }
}

// z and y of one x-plane per block (many small grids): forward, the plane's real rows two per complex transform
// along z (as kFftZ), their half spectra unpacked transposed (a line of K1 per z), then the y transforms; inverse
// the other way round.  The plane goes through global memory once each way instead of twice.
template<class T, bool INVERSE>
__global__ void __launch_bounds__(kBlock) kFftZY(const FftGroup* __restrict__ groups, int ngroups,
                                                 T* __restrict__ realAll, typename Vec2<T>::type* __restrict__ cplxAll,
                                                 const typename Vec2<T>::type* __restrict__ twAll, const int* skip,
                                                 long long* __restrict__ fixedAll, size_t fixedStride) {
    typedef typename Vec2<T>::type T2;
    extern __shared__ unsigned char smemRaw[];
    if (skip && *skip)
        return;
    const FftGroup& grp = groups[fftGroupOf(groups, ngroups, 3)];
    const PmeGrid& g = grp.g;
    const int K0 = g.K[0], K1 = g.K[1], K2 = g.K[2], Kh = g.K3h;
    const int plane = (int) blockIdx.x-grp.blk0[3];   // grid * K0 + x
    T* __restrict__ real = realAll+grp.realOff+(size_t) plane*K1*K2;
    T2* __restrict__ cplx = cplxAll+grp.cplxOff+(size_t) plane*K1*Kh;
    (void) K0;
    const int npairs = (K1+1)/2;
    const int S = max(npairs*K2, Kh*K1);
    T2* twY = reinterpret_cast<T2*>(smemRaw);
    T2* twZ = twY+K1;
    T2* bufA = twZ+K2;
    T2* bufB = bufA+S;
    loadTwiddles<T>(twY, twAll+grp.twOff+g.K[0], K1);
    loadTwiddles<T>(twZ, twAll+grp.twOff+g.K[0]+K1, K2);
    if (!INVERSE) {
        long long* fixed = fixedAll ? fixedAll+grp.realOff+(size_t) plane*K1*K2 : nullptr;
        for (int i = threadIdx.x; i < npairs*K2; i += blockDim.x) {
            const int l = i/K2, k = i-l*K2;
            const int ya = 2*l, yb = ya+1;
            T2 z;
            if (fixed) {
                z.x = takeFixed<T>(fixed+(size_t) ya*K2+k, fixedStride);
                z.y = (yb < K1 ? takeFixed<T>(fixed+(size_t) yb*K2+k, fixedStride) : T(0));
            }
            else {
                z.x = real[(size_t) ya*K2+k];
                z.y = (yb < K1 ? real[(size_t) yb*K2+k] : T(0));
            }
            bufA[l*K2+k] = z;
        }
        __syncthreads();
        T2* res = ldsFft<T>(bufA, bufB, twZ, g.plan[2], npairs, false);
        T2* other = (res == bufA ? bufB : bufA);
        for (int i = threadIdx.x; i < npairs*Kh; i += blockDim.x) {
            const int l = i/Kh, k = i-l*Kh;
            const T2 Z = res[l*K2+k], Zc = res[l*K2+(K2-k)%K2];
            T2 X, Y;
            X.x = T(0.5)*(Z.x+Zc.x);
            X.y = T(0.5)*(Z.y-Zc.y);
            Y.x = T(0.5)*(Z.y+Zc.y);
            Y.y = T(-0.5)*(Z.x-Zc.x);
            other[k*K1+2*l] = X;
            if (2*l+1 < K1)
                other[k*K1+2*l+1] = Y;
        }
        __syncthreads();
        T2* res2 = ldsFft<T>(other, res, twY, g.plan[1], Kh, false);
        for (int i = threadIdx.x; i < K1*Kh; i += blockDim.x) {
            const int y = i/Kh, z = i-y*Kh;
            cplx[(size_t) y*Kh+z] = res2[z*K1+y];
        }
    }
    else {
        for (int i = threadIdx.x; i < K1*Kh; i += blockDim.x) {
            const int y = i/Kh, z = i-y*Kh;
            bufA[z*K1+y] = cplx[(size_t) y*Kh+z];
        }
        __syncthreads();
        T2* res = ldsFft<T>(bufA, bufB, twY, g.plan[1], Kh, true);
        T2* other = (res == bufA ? bufB : bufA);
        for (int i = threadIdx.x; i < npairs*K2; i += blockDim.x) {
            const int l = i/K2, k = i-l*K2;
            const int ya = 2*l, yb = ya+1;
            const bool lower = k < Kh;
            const int kk = lower ? k : K2-k;
            T2 X = res[kk*K1+ya];
            T2 Y;
            Y.x = Y.y = T(0);
            if (yb < K1)
                Y = res[kk*K1+yb];
            if (kk == 0 || (K2%2 == 0 && kk == K2/2)) {
                X.y = T(0);
                Y.y = T(0);
            }
            T2 z;
            if (lower) {
                z.x = X.x-Y.y;
                z.y = X.y+Y.x;
            }
            else {
                z.x = X.x+Y.y;
                z.y = Y.x-X.y;
            }
            other[l*K2+k] = z;
        }
        __syncthreads();
        T2* res2 = ldsFft<T>(other, res, twZ, g.plan[2], npairs, true);
        for (int i = threadIdx.x; i < npairs*K2; i += blockDim.x) {
            const int l = i/K2, k = i-l*K2;
            const int ya = 2*l, yb = ya+1;
            const T2 z = res2[l*K2+k];
            real[(size_t) ya*K2+k] = z.x;
            if (yb < K1)
                real[(size_t) yb*K2+k] = z.y;
        }
    }
}

// x: forward, influence function (energy, strain), back; a tile of consecutive z for fixed y per block.
template<class T>
__global__ void __launch_bounds__(kBlock) kFftXConv(const FftGroup* __restrict__ groups, int ngroups,
                                                    typename Vec2<T>::type* __restrict__ cplxAll,
                                                    const T* __restrict__ inflAll,
                                                    const typename Vec2<T>::type* __restrict__ twAll,
                                                    BoxT<double> boxOne, const BoxT<double>* boxes, double beta,
                                                    int energy, int strain, double* __restrict__ partials,
                                                    const int* skip) {
    typedef typename Vec2<T>::type T2;
    extern __shared__ unsigned char smemRaw[];
    __shared__ double red[kBlock/64][10];
    if (skip && *skip)
        return;
    const FftGroup& grp = groups[fftGroupOf(groups, ngroups, 2)];
    const PmeGrid& g = grp.g;
    const int zt = grp.zx, blk = (int) blockIdx.x-grp.blk0[2];
    const int K = g.K[0], K2 = g.K[1], K3 = g.K[2], Kh = g.K3h;
    const int ntile = (Kh+zt-1)/zt;
    // stacked grids of one shape: block -> (grid, y, z tile)
    const int gi = blk/(K2*ntile), lb = blk-gi*(K2*ntile);
    const size_t goff = (size_t) grp.cplxOff+(size_t) gi*K*K2*Kh;
    typename Vec2<T>::type* __restrict__ cplx = cplxAll+goff;
    const T* __restrict__ infl = inflAll+goff;
    const BoxT<double>& box = boxes ? boxes[grp.boxOff+gi] : boxOne;
    const int y = lb/ntile, z0 = (lb%ntile)*zt;
    const int nl = min(zt, Kh-z0);
    T2* tw = reinterpret_cast<T2*>(smemRaw);
    T2* b0 = tw+K;
    T2* b1 = b0+zt*K;
    loadTwiddles<T>(tw, twAll+grp.twOff, K);
    for (int i = threadIdx.x; i < nl*K; i += blockDim.x) {
        const int xx = i/nl, l = i-xx*nl;
        b0[l*K+xx] = cplx[((size_t) xx*K2+y)*Kh+z0+l];
    }
    __syncthreads();
    T2* res = ldsFft<T>(b0, b1, tw, g.plan[0], nl, false);
    T2* other = (res == b0 ? b1 : b0);
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double fac = kPi*kPi/(beta*beta);
    for (int i = threadIdx.x; i < nl*K; i += blockDim.x) {
        const int l = i/K, m1 = i-l*K, m3 = z0+l;
        T2 F = res[l*K+m1];
        const T G = infl[((size_t) m1*K2+y)*Kh+m3];
        if (energy) {
            const double w = (m3 == 0 || (K3%2 == 0 && m3 == K3/2)) ? 1.0 : 2.0;
            const double em = 0.5*w*(double) G*((double) F.x*F.x+(double) F.y*F.y);
            acc[0] += em;
            if (strain && G != T(0)) {
                const int f1 = (m1 <= K/2 ? m1 : m1-K), f2 = (y <= K2/2 ? y : y-K2);
                double mv[3];
#pragma unroll
                for (int k = 0; k < 3; k++)
                    mv[k] = f1*box.recip[k][0]+f2*box.recip[k][1]+m3*box.recip[k][2];
                const double c2 = 2.0*(fac+1.0/(mv[0]*mv[0]+mv[1]*mv[1]+mv[2]*mv[2]));
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++)
                        acc[1+3*a+b] += em*((a == b ? -1.0 : 0.0)+c2*mv[a]*mv[b]);
            }
        }
        F.x *= G;
        F.y *= G;
        res[l*K+m1] = F;
    }
    __syncthreads();
    T2* res2 = ldsFft<T>(res, other, tw, g.plan[0], nl, true);
    for (int i = threadIdx.x; i < nl*K; i += blockDim.x) {
        const int xx = i/nl, l = i-xx*nl;
        cplx[((size_t) xx*K2+y)*Kh+z0+l] = res2[l*K+xx];
    }
    if (energy) {
        const int nq = strain ? 10 : 1;
        for (int q = 0; q < nq; q++) {
            const double v = groupSum<64>(acc[q]);
            if ((threadIdx.x & 63) == 0)
                red[threadIdx.x >> 6][q] = v;
        }
        __syncthreads();
        if (threadIdx.x < 10) {
            double v = 0;
            if (threadIdx.x < nq)
                for (int w = 0; w < kBlock/64; w++)
                    v += red[w][threadIdx.x];
            partials[((size_t) grp.partOff+blk)*10+threadIdx.x] = v;
        }
    }
}

// ---------------------------------------------------------------------------------------- gathering

constexpr int kGatherLanes = 8;

template<class T, int P, bool FINAL, bool BATCH>
__global__ void __launch_bounds__(kBlock) kGather(PmeSel ps, int n, const int* __restrict__ pmeAtom,
                                                  const int* __restrict__ baseArr, const T* __restrict__ spl,
                                                  const T* __restrict__ phiAll,
                                                  const typename Vec4<T>::type* __restrict__ dip,
                                                  const typename Vec4<T>::type* __restrict__ posq,
                                                  double* __restrict__ field, double* __restrict__ force,
                                                  const int* skip) {
    typedef typename Vec4<T>::type T4;
    if (skip && *skip)
        return;
    const int gid = blockIdx.x*blockDim.x+threadIdx.x;
    const int k = gid/kGatherLanes, lane = gid%kGatherLanes;
    if (k >= n)
        return;
    const int s = pmeAtom[k];
    const PmeSys* se = sysEntry<BATCH>(ps, BATCH ? ps.sysOf[s] : 0);
    const PmeGrid& g = BATCH ? se->g : ps.g;
    const BoxT<double>& box = BATCH ? se->box : ps.box;
    const T* __restrict__ phi = phiAll+(BATCH ? se->gridOff : 0);
    const T* o = spl+(size_t) k*9*P;
    const int B0 = baseArr[3*k], B1 = baseArr[3*k+1], B2 = baseArr[3*k+2];
    const int K0 = g.K[0], K1 = g.K[1], K2 = g.K[2];
    T t3[P], u3[P], v3[P];
    int g3[P];
#pragma unroll
    for (int j = 0; j < P; j++) {
        t3[j] = o[6*P+j];
        u3[j] = o[7*P+j];
        v3[j] = o[8*P+j];
        int z = B2-j;
        g3[j] = z < 0 ? z+K2 : z;
    }
    double f[3] = {0, 0, 0}, h[6] = {0, 0, 0, 0, 0, 0};
    for (int c = lane; c < P*P; c += kGatherLanes) {
        const int j1 = c/P, j2 = c-j1*P;
        int x0 = B0-j1, x1 = B1-j2;
        x0 += (x0 < 0 ? K0 : 0);
        x1 += (x1 < 0 ? K1 : 0);
        const T* row = phi+((size_t) x0*K1+x1)*K2;
        T ph[P];
        if (B2 >= P-1) {   // no wrap along z: the P points are contiguous and load as a few wide loads
            const T* run = row+B2-(P-1);
#pragma unroll
            for (int j = 0; j < P; j++)
                ph[P-1-j] = run[j];
        }
        else {
#pragma unroll
            for (int j3 = 0; j3 < P; j3++)
                ph[j3] = row[g3[j3]];
        }
        T st = 0, su = 0, sv = 0;
#pragma unroll
        for (int j3 = 0; j3 < P; j3++) {
            st += t3[j3]*ph[j3];
            su += u3[j3]*ph[j3];
            if (FINAL)
                sv += v3[j3]*ph[j3];
        }
        const T t1 = o[j1], u1 = o[P+j1], t2 = o[3*P+j2], u2 = o[4*P+j2];
        f[0] += (double) (u1*t2*st);
        f[1] += (double) (t1*u2*st);
        f[2] += (double) (t1*t2*su);
        if (FINAL) {
            const T v1 = o[2*P+j1], v2 = o[5*P+j2];
            h[0] += (double) (v1*t2*st);   // 00
            h[1] += (double) (t1*v2*st);   // 11
            h[2] += (double) (t1*t2*sv);   // 22
            h[3] += (double) (u1*u2*st);   // 01
            h[4] += (double) (u1*t2*su);   // 02
            h[5] += (double) (t1*u2*su);   // 12
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
        f[a] = groupSum<kGatherLanes>(f[a]);
    if (FINAL)
#pragma unroll
        for (int a = 0; a < 6; a++)
            h[a] = groupSum<kGatherLanes>(h[a]);
    if (lane != 0)
        return;
    double rk[3][3];
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
#pragma unroll
        for (int a = 0; a < 3; a++)
            rk[kk][a] = g.K[a]*box.recip[kk][a];
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
        field[3*s+kk] = -(rk[kk][0]*f[0]+rk[kk][1]*f[1]+rk[kk][2]*f[2]);
    if (FINAL) {
        const T4 d = dip[s];
        const double q = (double) posq[s].w;
        double e[3];
#pragma unroll
        for (int a = 0; a < 3; a++)
            e[a] = rk[0][a]*d.x+rk[1][a]*d.y+rk[2][a]*d.z;
        const double H[3][3] = {{h[0], h[3], h[4]}, {h[3], h[1], h[5]}, {h[4], h[5], h[2]}};
        double c[3];
#pragma unroll
        for (int a = 0; a < 3; a++)
            c[a] = q*f[a]+e[0]*H[a][0]+e[1]*H[a][1]+e[2]*H[a][2];
#pragma unroll
        for (int kk = 0; kk < 3; kk++)
            force[3*s+kk] = -(rk[kk][0]*c[0]+rk[kk][1]*c[1]+rk[kk][2]*c[2]);
    }
}

// Potential and field of the convolved grid phi at the atoms and, TAN, their tangents when every atom moves by g_s
// and the grid changes by phiDot: phi' = g . grad phi + phiDot, E' = -(grad grad phi) g - grad phiDot.
// out (FP64, sorted order): phi[n], E[3n] at offsets 0 and n, phi'[n], E'[3n] at offsets 5n and 6n.
template<class T, int P, bool TAN, bool BATCH>
__global__ void __launch_bounds__(kBlock) kGatherTangent(PmeSel ps, int n, const int* __restrict__ pmeAtom,
                                                         const int* __restrict__ baseArr, const T* __restrict__ spl,
                                                         const T* __restrict__ phiAll,
                                                         const T* __restrict__ phiDotAll,
                                                         const typename Vec4<T>::type* __restrict__ disp,
                                                         double* __restrict__ out) {
    typedef typename Vec4<T>::type T4;
    const int gid = blockIdx.x*blockDim.x+threadIdx.x;
    const int k = gid/kGatherLanes, lane = gid%kGatherLanes;
    if (k >= n)
        return;
    const int s = pmeAtom[k];
    const PmeSys* se = sysEntry<BATCH>(ps, BATCH ? ps.sysOf[s] : 0);
    const PmeGrid& g = BATCH ? se->g : ps.g;
    const BoxT<double>& box = BATCH ? se->box : ps.box;
    const T* __restrict__ phi = phiAll+(BATCH ? se->gridOff : 0);
    const T* __restrict__ phiDot = phiDotAll+(BATCH ? se->gridOff : 0);
    const T* o = spl+(size_t) k*9*P;
    const int B0 = baseArr[3*k], B1 = baseArr[3*k+1], B2 = baseArr[3*k+2];
    const int K0 = g.K[0], K1 = g.K[1], K2 = g.K[2];
    T t3[P], u3[P], v3[P];
    int g3[P];
#pragma unroll
    for (int j = 0; j < P; j++) {
        t3[j] = o[6*P+j];
        u3[j] = o[7*P+j];
        v3[j] = o[8*P+j];
        int z = B2-j;
        g3[j] = z < 0 ? z+K2 : z;
    }
    // pot, f[3], h[6] of phi; pot', f'[3] of phiDot
    double acc[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = lane; c < P*P; c += kGatherLanes) {
        const int j1 = c/P, j2 = c-j1*P;
        int x0 = B0-j1, x1 = B1-j2;
        x0 += (x0 < 0 ? K0 : 0);
        x1 += (x1 < 0 ? K1 : 0);
        const size_t rowOff = ((size_t) x0*K1+x1)*K2;
        T ph[P], pd[P];
        if (B2 >= P-1) {   // no wrap along z: contiguous points, a few wide loads (as in kGather)
            const size_t run = rowOff+B2-(P-1);
#pragma unroll
            for (int j = 0; j < P; j++) {
                ph[P-1-j] = phi[run+j];
                if (TAN)
                    pd[P-1-j] = phiDot[run+j];
            }
        }
        else {
#pragma unroll
            for (int j3 = 0; j3 < P; j3++) {
                ph[j3] = phi[rowOff+g3[j3]];
                if (TAN)
                    pd[j3] = phiDot[rowOff+g3[j3]];
            }
        }
        T st = 0, su = 0, sv = 0, dt = 0, du = 0;
#pragma unroll
        for (int j3 = 0; j3 < P; j3++) {
            st += t3[j3]*ph[j3];
            su += u3[j3]*ph[j3];
            if (TAN) {
                sv += v3[j3]*ph[j3];
                dt += t3[j3]*pd[j3];
                du += u3[j3]*pd[j3];
            }
        }
        const T t1 = o[j1], u1 = o[P+j1], v1 = o[2*P+j1], t2 = o[3*P+j2], u2 = o[4*P+j2], v2 = o[5*P+j2];
        acc[0] += (double) (t1*t2*st);
        acc[1] += (double) (u1*t2*st);
        acc[2] += (double) (t1*u2*st);
        acc[3] += (double) (t1*t2*su);
        if (TAN) {
            acc[4] += (double) (v1*t2*st);   // 00
            acc[5] += (double) (t1*v2*st);   // 11
            acc[6] += (double) (t1*t2*sv);   // 22
            acc[7] += (double) (u1*u2*st);   // 01
            acc[8] += (double) (u1*t2*su);   // 02
            acc[9] += (double) (t1*u2*su);   // 12
            acc[10] += (double) (t1*t2*dt);
            acc[11] += (double) (u1*t2*dt);
            acc[12] += (double) (t1*u2*dt);
            acc[13] += (double) (t1*t2*du);
        }
    }
#pragma unroll
    for (int a = 0; a < 14; a++)
        if (TAN || a < 4)
            acc[a] = groupSum<kGatherLanes>(acc[a]);
    if (lane != 0)
        return;
    double rk[3][3];
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
#pragma unroll
        for (int a = 0; a < 3; a++)
            rk[kk][a] = g.K[a]*box.recip[kk][a];
    out[s] = acc[0];
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
        out[n+3*s+kk] = -(rk[kk][0]*acc[1]+rk[kk][1]*acc[2]+rk[kk][2]*acc[3]);
    if (TAN) {
        const T4 gs = disp[s];
        double ga[3];
#pragma unroll
        for (int a = 0; a < 3; a++)
            ga[a] = rk[0][a]*gs.x+rk[1][a]*gs.y+rk[2][a]*gs.z;
        const double H[3][3] = {{acc[4], acc[7], acc[8]}, {acc[7], acc[5], acc[9]}, {acc[8], acc[9], acc[6]}};
        out[5*n+s] = ga[0]*acc[1]+ga[1]*acc[2]+ga[2]*acc[3]+acc[10];
        double c[3];
#pragma unroll
        for (int a = 0; a < 3; a++)
            c[a] = H[a][0]*ga[0]+H[a][1]*ga[1]+H[a][2]*ga[2]+acc[11+a];
#pragma unroll
        for (int kk = 0; kk < 3; kk++)
            out[6*n+3*s+kk] = -(rk[kk][0]*c[0]+rk[kk][1]*c[1]+rk[kk][2]*c[2]);
    }
}

// ---------------------------------------------------------------------------------------- setup

template<class T>
__global__ void kInfluence(PmeGrid g, const double* bmod, BoxT<double> box, double beta, double volume, T* infl) {
    const int K1 = g.K[0], K2 = g.K[1], Kh = g.K3h;
    const size_t total = (size_t) K1*K2*Kh;
    const size_t idx = (size_t) blockIdx.x*blockDim.x+threadIdx.x;
    if (idx >= total)
        return;
    const int m3 = idx%Kh, m2 = (idx/Kh)%K2, m1 = idx/((size_t) Kh*K2);
    const int f1 = (m1 <= K1/2 ? m1 : m1-K1), f2 = (m2 <= K2/2 ? m2 : m2-K2);
    double val = 0;
    if (f1 != 0 || f2 != 0 || m3 != 0) {
        double mv[3];
        for (int k = 0; k < 3; k++)
            mv[k] = f1*box.recip[k][0]+f2*box.recip[k][1]+m3*box.recip[k][2];
        const double m2s = mv[0]*mv[0]+mv[1]*mv[1]+mv[2]*mv[2];
        const double b = bmod[m1]*bmod[K1+m2]*bmod[K1+K2+m3];
        val = exp(-kPi*kPi/(beta*beta)*m2s)/(kPi*volume*m2s)/b;
    }
    infl[idx] = (T) val;
}

template<class T>
__global__ void kTwiddles(PmeGrid g, typename Vec2<T>::type* tw) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    const int total = g.K[0]+g.K[1]+g.K[2];
    if (i >= total)
        return;
    const int a = (i < g.K[0] ? 0 : (i < g.K[0]+g.K[1] ? 1 : 2));
    const int j = i-(a == 0 ? 0 : (a == 1 ? g.K[0] : g.K[0]+g.K[1]));
    double sn, cs;
    sincospi(-2.0*j/g.K[a], &sn, &cs);
    typename Vec2<T>::type v;
    v.x = (T) cs;
    v.y = (T) sn;
    tw[i] = v;
}

template<class T>
__global__ void kZero(T* p, size_t count, const int* skip) {
    if (skip && *skip)
        return;
    for (size_t i = (size_t) blockIdx.x*blockDim.x+threadIdx.x; i < count; i += (size_t) gridDim.x*blockDim.x)
        p[i] = T(0);
}

inline int nblk(long long threads, int bs) { return (int) ((threads+bs-1)/bs); }

template<class T>
int linesPerBlock(int K, long long lines, int ldsBytes) {
    // LDS: twiddles + two line buffers within ldsBytes; and enough blocks to fill 256 CUs twice over.  More
    // lines per block keep more of the 256 threads busy in each radix stage.
    const int elems = (int) (ldsBytes/sizeof(typename Vec2<T>::type));
    const int byLds = std::max(1, std::min(64, (elems-K)/(2*K)));
    const int byFill = (int) std::max(1LL, lines/512);
    return std::min(byLds, byFill);
}

// The y and x passes tile a row of K3h half-spectrum lines and read each tile as runs of consecutive z: up to
// 64 KB of LDS (less the x pass's static reduction buffer) gives runs of whole 128-byte lines where the grid
// allows (16 complex floats or 8 complex doubles at K = 216, against 6 and 3 within 24 KB), and the tiles are
// balanced over the row, with no short tail tile.
template<class T>
int tileLines(int K, int Kh, long long lines) {
    const int L = linesPerBlock<T>(K, lines, 64*1024-512);
    const int ntile = (Kh+L-1)/L;
    return (Kh+ntile-1)/ntile;
}

}   // namespace

// Runs fn with the PME order and the batch flag as compile-time constants.
template<class Fn>
void dispatchOrder(int order, bool batch, Fn fn) {
    auto withBatch = [&](auto P) {
        if (batch)
            fn(P, std::true_type{});
        else
            fn(P, std::false_type{});
    };
    switch (order) {
    case 4:
        withBatch(std::integral_constant<int, 4>{});
        break;
    case 5:
        withBatch(std::integral_constant<int, 5>{});
        break;
    case 6:
        withBatch(std::integral_constant<int, 6>{});
        break;
    case 7:
        withBatch(std::integral_constant<int, 7>{});
        break;
    case 8:
        withBatch(std::integral_constant<int, 8>{});
        break;
    default:
        throw std::runtime_error("pgm_hip: PME order must be 4..8");
    }
}

template<class T>
void launchPmeBin(hipStream_t st, int n, const double4* posD, PmeSel ps, int* brickOf, int* slot, int* brickCount) {
    if (ps.table)
        kPmeBin<T, true><<<nblk(n, 256), 256, 0, st>>>(n, posD, ps, brickOf, slot, brickCount);
    else
        kPmeBin<T, false><<<nblk(n, 256), 256, 0, st>>>(n, posD, ps, brickOf, slot, brickCount);
}

template<class T>
void launchPmePlace(hipStream_t st, int n, const double4* posD, PmeSel ps, const int* brickOf, const int* slot,
                    const int* brickStart, int* pmeAtom, int* base, T* spl) {
    dispatchOrder(ps.order, ps.table != nullptr, [&](auto P, auto B) {
        kPmePlace<T, P(), B()><<<nblk(n, 128), 128, 0, st>>>(n, posD, ps, brickOf, slot, brickStart, pmeAtom, base,
                                                             spl);
    });
}

template<class T>
void launchSpread(hipStream_t st, PmeSel ps, const int* brickStart, const int* pmeAtom, const int* base, const T* spl,
                  const typename Vec4<T>::type* dip, const typename Vec4<T>::type* posq, int withQ, T* grid,
                  long long* fixed, const int* skip, const typename Vec4<T>::type* dipDot,
                  const typename Vec4<T>::type* disp) {
    const int E = ps.brick+ps.order-1;
    const size_t lds = (size_t) E*E*E*sizeof(double)*(dipDot ? 1 : GridFixed<T>::words);
    dispatchOrder(ps.order, ps.table != nullptr, [&](auto P, auto B) {
        if (dipDot)
            kSpread<T, P(), true, B()><<<ps.nbrick, kBlock, lds, st>>>(ps, brickStart, pmeAtom, base, spl, dip, posq,
                                                                       withQ, grid, nullptr, skip, dipDot, disp);
        else
            kSpread<T, P(), false, B()><<<ps.nbrick, kBlock, lds, st>>>(ps, brickStart, pmeAtom, base, spl, dip,
                                                                        posq, withQ, nullptr, fixed, skip, dip, dip);
    });
}

template<class T>
void fftGroupShape(FftGroup& G, int blocks[4], size_t lds[4]) {
    typedef typename Vec2<T>::type T2;
    const PmeGrid& g = G.g;
    const int ngrid = G.ngrid;
    const int nlines = ngrid*g.K[0]*g.K[1];
    // z: pairs of real lines; y: the stacked grids are one grid of ngrid K0 x-planes; x: lines of K0
    G.pz = linesPerBlock<T>(g.K[2], (nlines+1)/2, 24*1024);   // contiguous lines: small tiles, more blocks
    blocks[0] = nblk((nlines+1)/2, G.pz);
    lds[0] = (size_t) (g.K[2]+2*G.pz*g.K[2])*sizeof(T2);
    G.zy = tileLines<T>(g.K[1], g.K3h, (long long) ngrid*g.K[0]*g.K3h);
    blocks[1] = ngrid*g.K[0]*((g.K3h+G.zy-1)/G.zy);
    lds[1] = (size_t) (g.K[1]+2*G.zy*g.K[1])*sizeof(T2);
    G.zx = tileLines<T>(g.K[0], g.K3h, (long long) ngrid*g.K[1]*g.K3h);
    blocks[2] = ngrid*g.K[1]*((g.K3h+G.zx-1)/G.zx);
    lds[2] = (size_t) (g.K[0]+2*G.zx*g.K[0])*sizeof(T2);
    // fused z+y: one block per x-plane; twiddles of y and z and two plane buffers
    const int S = std::max(((g.K[1]+1)/2)*g.K[2], g.K3h*g.K[1]);
    blocks[3] = ngrid*g.K[0];
    lds[3] = (size_t) (g.K[1]+g.K[2]+2*S)*sizeof(T2);
}

template<class T>
void launchPmeConvolve(hipStream_t st, const FftLaunch& L, T* real, typename Vec2<T>::type* cplx, const T* influence,
                       const typename Vec2<T>::type* tw, BoxT<double> box, const BoxT<double>* boxes, double beta,
                       int energy, int strain, double* partials, const int* skip, long long* fixed,
                       size_t fixedStride) {
    const FftGroup* G = L.groups;
    const int ng = L.ngroups;
    if (L.zyFused)
        kFftZY<T, false><<<L.blocks[3], kBlock, L.lds[3], st>>>(G, ng, real, cplx, tw, skip, fixed, fixedStride);
    else {
        kFftZ<T, false><<<L.blocks[0], kBlock, L.lds[0], st>>>(G, ng, real, cplx, tw, skip, fixed, fixedStride);
        kFftY<T, false><<<L.blocks[1], kBlock, L.lds[1], st>>>(G, ng, cplx, tw, skip);
    }
    kFftXConv<T><<<L.blocks[2], kBlock, L.lds[2], st>>>(G, ng, cplx, influence, tw, box, boxes, beta, energy, strain,
                                                        partials, skip);
    if (L.zyFused)
        kFftZY<T, true><<<L.blocks[3], kBlock, L.lds[3], st>>>(G, ng, real, cplx, tw, skip, nullptr, 0);
    else {
        kFftY<T, true><<<L.blocks[1], kBlock, L.lds[1], st>>>(G, ng, cplx, tw, skip);
        kFftZ<T, true><<<L.blocks[0], kBlock, L.lds[0], st>>>(G, ng, real, cplx, tw, skip, nullptr, 0);
    }
}

template<class T>
int convolveBlocksPerGrid(PmeGrid g, int ngrid) {
    const int zx = tileLines<T>(g.K[0], g.K3h, (long long) ngrid*g.K[1]*g.K3h);
    return g.K[1]*((g.K3h+zx-1)/zx);
}

template<class T>
void launchGatherTangent(hipStream_t st, PmeSel ps, int n, const int* pmeAtom, const int* base, const T* spl,
                         const T* phi, const T* phiDot, const typename Vec4<T>::type* disp, double* out) {
    const int blocks = nblk((long long) n*kGatherLanes, kBlock);
    dispatchOrder(ps.order, ps.table != nullptr, [&](auto P, auto B) {
        if (phiDot)
            kGatherTangent<T, P(), true, B()><<<blocks, kBlock, 0, st>>>(ps, n, pmeAtom, base, spl, phi, phiDot, disp,
                                                                         out);
        else
            kGatherTangent<T, P(), false, B()><<<blocks, kBlock, 0, st>>>(ps, n, pmeAtom, base, spl, phi, phi, disp,
                                                                          out);
    });
}

template<class T>
void launchGather(hipStream_t st, PmeSel ps, int n, const int* pmeAtom, const int* base, const T* spl, const T* phi,
                  const typename Vec4<T>::type* dip, const typename Vec4<T>::type* posq, int final_, double* field,
                  double* force, const int* skip) {
    const int blocks = nblk((long long) n*kGatherLanes, kBlock);
    dispatchOrder(ps.order, ps.table != nullptr, [&](auto P, auto B) {
        if (final_)
            kGather<T, P(), true, B()><<<blocks, kBlock, 0, st>>>(ps, n, pmeAtom, base, spl, phi, dip, posq, field,
                                                                  force, skip);
        else
            kGather<T, P(), false, B()><<<blocks, kBlock, 0, st>>>(ps, n, pmeAtom, base, spl, phi, dip, posq, field,
                                                                   force, skip);
    });
}

template<class T>
void launchInfluence(hipStream_t st, PmeGrid g, const double* bmod, BoxT<double> box, double beta, double volume,
                     T* influence) {
    const size_t total = (size_t) g.K[0]*g.K[1]*g.K3h;
    kInfluence<T><<<nblk((long long) total, 256), 256, 0, st>>>(g, bmod, box, beta, volume, influence);
}

template<class T>
void launchTwiddles(hipStream_t st, PmeGrid g, typename Vec2<T>::type* tw) {
    kTwiddles<T><<<nblk(g.K[0]+g.K[1]+g.K[2], 256), 256, 0, st>>>(g, tw);
}

template<class T>
void launchZero(hipStream_t st, T* p, size_t count, const int* skip) {
    const int blocks = (int) std::min<size_t>(2048, (count+255)/256);
    kZero<T><<<std::max(1, blocks), 256, 0, st>>>(p, count, skip);
}

#define PGM_INSTANTIATE_PME(T)                                                                                          \
    template void launchPmeBin<T>(hipStream_t, int, const double4*, PmeSel, int*, int*, int*);                         \
    template void launchPmePlace<T>(hipStream_t, int, const double4*, PmeSel, const int*, const int*, const int*,      \
                                    int*, int*, T*);                                                                   \
    template void launchSpread<T>(hipStream_t, PmeSel, const int*, const int*, const int*, const T*,                   \
                                  const Vec4<T>::type*, const Vec4<T>::type*, int, T*, long long*, const int*,         \
                                  const Vec4<T>::type*, const Vec4<T>::type*);                                         \
    template void launchGatherTangent<T>(hipStream_t, PmeSel, int, const int*, const int*, const T*, const T*,         \
                                         const T*, const Vec4<T>::type*, double*);                                     \
    template void fftGroupShape<T>(FftGroup&, int*, size_t*);                                                       \
    template void launchPmeConvolve<T>(hipStream_t, const FftLaunch&, T*, Vec2<T>::type*, const T*,                    \
                                       const Vec2<T>::type*, BoxT<double>, const BoxT<double>*, double, int, int,     \
                                       double*, const int*, long long*, size_t);                                      \
    template void launchGather<T>(hipStream_t, PmeSel, int, const int*, const int*, const T*, const T*,                \
                                  const Vec4<T>::type*, const Vec4<T>::type*, int, double*, double*, const int*);      \
    template void launchInfluence<T>(hipStream_t, PmeGrid, const double*, BoxT<double>, double, double, T*);           \
    template void launchTwiddles<T>(hipStream_t, PmeGrid, Vec2<T>::type*);                                             \
    template void launchZero<T>(hipStream_t, T*, size_t, const int*);                                                 \
    template int convolveBlocksPerGrid<T>(PmeGrid, int);
PGM_INSTANTIATE_PME(float)
PGM_INSTANTIATE_PME(double)
template void launchZero<int>(hipStream_t, int*, size_t, const int*);
template void launchZero<unsigned>(hipStream_t, unsigned*, size_t, const int*);
template void launchZero<long long>(hipStream_t, long long*, size_t, const int*);

}   // namespace pgmhip
