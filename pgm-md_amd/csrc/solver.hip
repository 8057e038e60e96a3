// Induced-dipole solver and the per-atom stages around it, FP64 vectors in the sorted order:
//   guess       mu4 / mu3 / last-solution predictor from the history (global order) or alpha b
//               (core/pgm_core.cpp:307-328; initialGuess, core/cuda/pgm_cuda.cu:1318-1335);
//   CG          Jacobi-preconditioned CG with the flexible Polak-Ribiere beta, err = max|alpha r| / mean|alpha b|
//               (core/pgm_core.cpp:330-385).  Every reduction is written as per-block partials and reduced
//               again, identically, by every block of the consuming kernel: deterministic, no atomics;
//   post        peek, history push, total dipoles (core/pgm_core.cpp:386-397, 413-416);
//   final       total field, self/polarization sums, covalent chain rule, forces out
//               (core/pgm_core.cpp:488-534).
#include "kernels.hpp"

#include <stdexcept>

#include <algorithm>

namespace pgmhip {

namespace {

constexpr int kCgBlock = 256;

inline int nblk(long long threads, int bs) { return (int) ((threads+bs-1)/bs); }

// Sum (or max) of one partials region of nb entries, computed redundantly by every block; valid in all threads.
template<bool MAX>
__device__ __forceinline__ double reduceRegion(const double* reg, int nb, double* smem) {
    double v = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x)
        v = MAX ? fmax(v, reg[k]) : v+reg[k];
    v = MAX ? blockMax(v, smem) : blockSum(v, smem);
    if (threadIdx.x == 0)
        smem[32] = v;
    __syncthreads();
    const double r = smem[32];
    __syncthreads();
    return r;
}

__global__ void kGuess(int n, const int* order, const int* gidx, const double* hist, size_t stride, GuessCoef gc,
                       const double* alpha, const double* b, double* x) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= 3*n)
        return;
    const int s = k/3, c = k-3*s;
    const double a = alpha[s];
    if (gc.nc == 0) {
        x[k] = a*b[k];
        return;
    }
    const size_t g = 3*(size_t) gidx[order[s]]+c;
    double v = 0;
    for (int h = 0; h < gc.nc; h++)
        v += gc.c[h]*hist[gc.slot[h]*stride+g];
    x[k] = (a > 0 ? v : 0.0);
}

__global__ void kCombineB(int n, const double* ed, const double* er, const double* perm, double cself, double* b) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k < 3*n)
        b[k] = ed[k]+er[k]+cself*perm[k];
}

// The pair vector type of a 3-vector per atom, w = rgw[s] (the radius parameter, read by the FP32 field pass
// that evaluates its kernels on the fly) or 0.
template<class T>
__global__ void kToVec4(int n, const double* v, typename Vec4<T>::type* out, const T* rgw) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    typename Vec4<T>::type o;
    o.x = (T) v[3*s];
    o.y = (T) v[3*s+1];
    o.z = (T) v[3*s+2];
    o.w = rgw ? rgw[s] : T(0);
    out[s] = o;
}

// r = b - A x (polarizable rows); partials of sum |alpha b| and max |alpha r|.  Jacobi (not pc.pre): z = p =
// alpha/(1 - alpha cj) r and the partials of r.z; with the short-range preconditioner kPrecond<true> does that.
template<class T>
__global__ void __launch_bounds__(kCgBlock) kCgInit(int n, const double* alpha, const double* b, const double* x,
                                                    const double* td, const double* tr, double cself, double* r,
                                                    double* z, double* p, typename Vec4<T>::type* pT,
                                                    double* partials, Precond pc, const T* rgw) {
    __shared__ double smem[40];
    double nrm = 0, rz = 0, mx = 0;
    for (int s = blockIdx.x*blockDim.x+threadIdx.x; s < n; s += gridDim.x*blockDim.x) {
        const double a = alpha[s];
        const double f = a/(1.0-a*pc.cj);
        double pv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int k = 3*s+c;
            const double ax = (a > 0 ? x[k]/a-(td[k]+tr[k]+cself*x[k]) : 0.0);
            const double rv = (a > 0 ? b[k]-ax : 0.0);
            const double zv = f*rv;
            r[k] = rv;
            nrm += fabs(a*b[k]);
            mx = fmax(mx, fabs(a*rv));   // the tolerance is on |alpha r| with every preconditioner
            if (!pc.pre) {
                z[k] = zv;
                p[k] = zv;
                rz += rv*zv;
            }
            pv[c] = zv;
        }
        if (!pc.pre) {
            typename Vec4<T>::type o;
            o.x = (T) pv[0];
            o.y = (T) pv[1];
            o.z = (T) pv[2];
            o.w = rgw ? rgw[s] : T(0);
            pT[s] = o;
        }
    }
    nrm = blockSum(nrm, smem);
    rz = blockSum(rz, smem);
    mx = blockMax(mx, smem);
    if (threadIdx.x == 0) {
        partials[kPNorm*kCgMaxBlocks+blockIdx.x] = nrm;
        if (!pc.pre)
            partials[kPRz0*kCgMaxBlocks+blockIdx.x] = rz;
        partials[kPMax*kCgMaxBlocks+blockIdx.x] = mx;
    }
}

__global__ void __launch_bounds__(kCgBlock) kCgInitCheck(int n, int nb, double tol, int maxIter,
                                                         const double* partials, CgState* cg, int* done) {
    __shared__ double smem[40];
    const double nrm = reduceRegion<false>(partials+kPNorm*kCgMaxBlocks, nb, smem)/(3.0*n)+1e-300;
    const double mx = reduceRegion<true>(partials+kPMax*kCgMaxBlocks, nb, smem);
    if (threadIdx.x == 0) {
        const double err = mx/nrm;
        cg->norm = nrm;
        cg->err = err;
        cg->iterations = 0;
        done[0] = (err > tol && 0 < maxIter) ? 0 : 1;
    }
}

__global__ void __launch_bounds__(kCgBlock) kCgAp(int n, const double* alpha, const double* p, const double* td,
                                                  const double* tr, double cself, double* ap, double* partials,
                                                  const int* skip) {
    __shared__ double smem[40];
    if (*skip)
        return;
    double pap = 0;
    for (int k = blockIdx.x*blockDim.x+threadIdx.x; k < 3*n; k += gridDim.x*blockDim.x) {
        const double a = alpha[k/3];
        const double v = (a > 0 ? p[k]/a-(td[k]+tr[k]+cself*p[k]) : 0.0);
        ap[k] = v;
        pap += p[k]*v;
    }
    pap = blockSum(pap, smem);
    if (threadIdx.x == 0)
        partials[kPpAp*kCgMaxBlocks+blockIdx.x] = pap;
}

// x += step p, r' = r - step Ap; partials of max |alpha r'| and, Jacobi (not pc.pre), with z' = alpha/(1 - alpha
// cj) r', of z'.(r' - r) and r'.z' (with the short-range preconditioner kPrecond<false> does those).
__global__ void __launch_bounds__(kCgBlock) kCgUpdate(int n, int it, const double* alpha, const double* p,
                                                      const double* ap, double* x, double* r, double* z,
                                                      double* partials, const int* skip, Precond pc) {
    __shared__ double smem[40];
    if (*skip)
        return;
    const int nb = gridDim.x;
    const double pap = reduceRegion<false>(partials+kPpAp*kCgMaxBlocks, nb, smem);
    const double rz = reduceRegion<false>(partials+((it & 1) ? kPRz1 : kPRz0)*kCgMaxBlocks, nb, smem);
    const double step = rz/fmax(pap, 1e-300);
    double bnum = 0, rzn = 0, mx = 0;
    for (int k = blockIdx.x*blockDim.x+threadIdx.x; k < 3*n; k += gridDim.x*blockDim.x) {
        const double a = alpha[k/3];
        x[k] += step*p[k];
        const double rold = r[k];
        const double rn = rold-step*ap[k];
        mx = fmax(mx, fabs(a*rn));
        r[k] = rn;
        if (!pc.pre) {
            const double zn = a/(1.0-a*pc.cj)*rn;
            bnum += zn*(rn-rold);
            rzn += rn*zn;
            z[k] = zn;
        }
    }
    bnum = blockSum(bnum, smem);
    rzn = blockSum(rzn, smem);
    mx = blockMax(mx, smem);
    if (threadIdx.x == 0) {
        if (!pc.pre) {
            partials[kPBnum*kCgMaxBlocks+blockIdx.x] = bnum;
            partials[((it & 1) ? kPRz0 : kPRz1)*kCgMaxBlocks+blockIdx.x] = rzn;
        }
        partials[kPMax*kCgMaxBlocks+blockIdx.x] = mx;
    }
}

// The short-range preconditioner, z_s = alpha_s (r_s + sum_j T_sj alpha_j r_j), T_sj u = -u G1 + (u.x) x G2 over
// the atom's short list (partners with alpha_j > 0; *over: z_s = alpha_s r_s everywhere), FP64 sums of the
// pair terms from the stored kernels and the positions of the pair type (core/cuda/pgm_cuda.cu:1533-1570).
// INIT (after kCgInit): p = z and its pair-type copy, partials of r.z; else (after kCgUpdate of iteration it):
// partials of -step z.Ap (= z.(r' - r)) and r.z.  One thread per atom, the CG kernels' grid.
template<class T, bool INIT, bool BATCH>
__global__ void __launch_bounds__(kCgBlock) kPrecond(int n, int it, const double* alpha, const double* r,
                                                     const double* ap, const typename Vec4<T>::type* pos,
                                                     BoxSel<T> bs, const int* srCount, const int* srJ,
                                                     const typename Vec2<T>::type* srG, const int* over, double* z,
                                                     double* p, typename Vec4<T>::type* pT, double* partials,
                                                     const int* skip, const T* rgw) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    __shared__ double smem[40];
    if (!INIT && *skip)
        return;
    const int nb = gridDim.x;
    double step = 0;
    if (!INIT) {
        const double pap = reduceRegion<false>(partials+kPpAp*kCgMaxBlocks, nb, smem);
        const double rz = reduceRegion<false>(partials+((it & 1) ? kPRz1 : kPRz0)*kCgMaxBlocks, nb, smem);
        step = rz/fmax(pap, 1e-300);
    }
    const bool useT = (*over == 0);
    double rz = 0, bnum = 0;
    for (int s = blockIdx.x*blockDim.x+threadIdx.x; s < n; s += gridDim.x*blockDim.x) {
        const double a = alpha[s];
        double e[3] = {0, 0, 0};
        if (useT && a > 0) {
            const BoxT<T> box = selectBox<BATCH>(bs, s);
            const T4 pi = pos[s];
            const int cnt = srCount[s];
            for (int q = 0; q < cnt; q++) {
                const int v = srJ[(size_t) s*kPreCap+q], j = entryAtom(v, bs.ext);
                const double aj = alpha[j];
                if (aj == 0)
                    continue;
                const T2 g = srG[(size_t) s*kPreCap+q];
                const T4 pj = pos[j];
                T x = pi.x-pj.x, y = pi.y-pj.y, zz = pi.z-pj.z;
                pairImage(x, y, zz, box, v, bs.ext);
                const double xd[3] = {(double) x, (double) y, (double) zz};
                const double u[3] = {aj*r[3*j], aj*r[3*j+1], aj*r[3*j+2]};
                const double ux = u[0]*xd[0]+u[1]*xd[1]+u[2]*xd[2];
#pragma unroll
                for (int c = 0; c < 3; c++)
                    e[c] += -u[c]*(double) g.x+ux*xd[c]*(double) g.y;
            }
        }
        double zv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int k = 3*s+c;
            zv[c] = a*(r[k]+e[c]);
            z[k] = zv[c];
            rz += r[k]*zv[c];
            if (INIT)
                p[k] = zv[c];
            else
                bnum -= step*zv[c]*ap[k];
        }
        if (INIT) {
            T4 o;
            o.x = (T) zv[0];
            o.y = (T) zv[1];
            o.z = (T) zv[2];
            o.w = rgw ? rgw[s] : T(0);
            pT[s] = o;
        }
    }
    rz = blockSum(rz, smem);
    if (!INIT)
        bnum = blockSum(bnum, smem);
    if (threadIdx.x == 0) {
        if (INIT)
            partials[kPRz0*kCgMaxBlocks+blockIdx.x] = rz;
        else {
            partials[kPBnum*kCgMaxBlocks+blockIdx.x] = bnum;
            partials[((it & 1) ? kPRz0 : kPRz1)*kCgMaxBlocks+blockIdx.x] = rz;
        }
    }
}

template<class T>
__global__ void __launch_bounds__(kCgBlock) kCgDirection(int n, int it, double tol, int maxIter, const double* z,
                                                         double* p, typename Vec4<T>::type* pT,
                                                         const double* partials, CgState* cg, int* done,
                                                         const T* rgw) {
    __shared__ double smem[40];
    if (done[it]) {
        // finished earlier: the flag of the next iteration (enqueued ahead) must say so too
        if (blockIdx.x == 0 && threadIdx.x == 0)
            done[it+1] = 1;
        return;
    }
    const int nb = gridDim.x;
    const double bnum = reduceRegion<false>(partials+kPBnum*kCgMaxBlocks, nb, smem);
    const double rzOld = reduceRegion<false>(partials+((it & 1) ? kPRz1 : kPRz0)*kCgMaxBlocks, nb, smem);
    const double bc = bnum/fmax(rzOld, 1e-300);
    for (int s = blockIdx.x*blockDim.x+threadIdx.x; s < n; s += gridDim.x*blockDim.x) {
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int k = 3*s+c;
            v[c] = z[k]+bc*p[k];
            p[k] = v[c];
        }
        typename Vec4<T>::type o;
        o.x = (T) v[0];
        o.y = (T) v[1];
        o.z = (T) v[2];
        o.w = rgw ? rgw[s] : T(0);
        pT[s] = o;
    }
    if (blockIdx.x == 0) {
        const double mx = reduceRegion<true>(partials+kPMax*kCgMaxBlocks, nb, smem);
        if (threadIdx.x == 0) {
            const double err = mx/cg->norm;
            cg->err = err;
            cg->iterations = it+1;
            done[it+1] = (err > tol && it+1 < maxIter) ? 0 : 1;
        }
    }
}

template<class T>
__global__ void kPost(int n, double peek, const double* alpha, const double* r, double* x, const double* perm,
                      const int* order, const int* gidx, double* hist, double* d, const double* rgS,
                      typename Vec4<T>::type* dT) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const double a = alpha[s];
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int k = 3*s+c;
        double xv = x[k];
        if (peek != 0)
            xv += peek*a*r[k];
        x[k] = xv;
        if (hist)
            hist[3*(size_t) gidx[order[s]]+c] = xv;
        v[c] = perm[k]+xv;
        d[k] = v[c];
    }
    typename Vec4<T>::type o;
    o.x = (T) v[0];
    o.y = (T) v[1];
    o.z = (T) v[2];
    o.w = (T) rgS[s];   // the radius parameter rides along for the pair passes' partner gathers
    dT[s] = o;
}

__global__ void __launch_bounds__(kCgBlock) kFinalCombine(int n, const double* d, const double* mu,
                                                          const double* alpha, const double* edir,
                                                          const double* erec, double cself, double* etot,
                                                          double* perAtom) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const double a = alpha[s];
    double dv[3], er[3], d2 = 0, pol = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        dv[c] = d[3*s+c];
        er[c] = erec[3*s+c];
        etot[3*s+c] = edir[3*s+c]+er[c]+cself*dv[c];
        d2 += dv[c]*dv[c];
        if (a > 0)
            pol += mu[3*s+c]*mu[3*s+c]/(2.0*a);
    }
    double* o = perAtom+11*(size_t) s;
    o[0] = d2;
    o[1] = pol;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
            o[2+3*p+q] = er[p]*dv[q];
}

// dE/dr_partner = c (g - (g.u) u)/|v|, g = -E_atom: +t on the atom, -t on the partner, strain t (x) v
// (core/pgm_core.cpp:512-529).  One thread per atom sums the entries it carries and those it is partner of.
__device__ __forceinline__ void covalentTerm(int m, const int* covAtomOf, const int* covPartner,
                                             const double* covStrength, const double* pos, const int* rank,
                                             const double* etot, const BoxT<double>& box, double t[3], double v[3]) {
    const int i = covAtomOf[m], j = covPartner[m];
    double x = pos[3*j]-pos[3*i], y = pos[3*j+1]-pos[3*i+1], z = pos[3*j+2]-pos[3*i+2];
    minImage(x, y, z, box);
    const double len = sqrt(x*x+y*y+z*z);
    const double u[3] = {x/len, y/len, z/len};
    const int si = rank[i];
    const double g[3] = {-etot[3*si], -etot[3*si+1], -etot[3*si+2]};
    const double gu = g[0]*u[0]+g[1]*u[1]+g[2]*u[2];
    const double c = covStrength[m]/len;
#pragma unroll
    for (int k = 0; k < 3; k++)
        t[k] = c*(g[k]-gu*u[k]);
    v[0] = x;
    v[1] = y;
    v[2] = z;
}

// Per atom i (caller's order): the chain-rule force of the entries it carries (+t) and of those it is the
// partner of (-t), and the strain of the entries it carries.  Atom and partner share a system.
template<bool BATCH>
__global__ void __launch_bounds__(kCgBlock) kCovalent(int n, const int* rank, const double* pos, const int* covStart,
                                                      const int* covPartner, const double* covStrength,
                                                      const int* revStart, const int* revEntry,
                                                      const int* covAtomOf, const double* etot, BoxSel<double> bs,
                                                      double* fcov, double* perAtom) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n)
        return;
    const BoxT<double> box = selectBox<BATCH>(bs, i);
    double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double f[3] = {0, 0, 0}, t[3], v[3];
    for (int m = covStart[i]; m < covStart[i+1]; m++) {
        covalentTerm(m, covAtomOf, covPartner, covStrength, pos, rank, etot, box, t, v);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            f[k] += t[k];
#pragma unroll
            for (int b = 0; b < 3; b++)
                S[3*k+b] += t[k]*v[b];
        }
    }
    for (int e = revStart[i]; e < revStart[i+1]; e++) {
        covalentTerm(revEntry[e], covAtomOf, covPartner, covStrength, pos, rank, etot, box, t, v);
#pragma unroll
        for (int k = 0; k < 3; k++)
            f[k] -= t[k];
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        fcov[3*i+k] = f[k];
#pragma unroll
    for (int q = 0; q < 9; q++)
        perAtom[9*(size_t) i+q] = S[q];
}

__global__ void kForcesOut(int n, const int* rank, const double* fpair, const double* frec, const double* fcov,
                           double ke, double* fOut, long long* fixedOut, int stride, double fixedScale,
                           const int* guard) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n || (guard && *guard == 0))
        return;
    const int s = rank[i];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double f = ke*(fpair[3*s+k]+frec[3*s+k]+fcov[3*i+k]);
        if (fOut)
            fOut[3*i+k] += f;
        if (fixedOut)
            atomicAdd(reinterpret_cast<unsigned long long*>(fixedOut+i+(size_t) k*stride),
                      (unsigned long long) llrint(f*fixedScale));
    }
}

__global__ void __launch_bounds__(256) kReducePartials(const double* partials, int nblk, int width, double* out) {
    __shared__ double smem[40];
    for (int q = 0; q < width; q++) {
        double v = 0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x)
            v += partials[(size_t) b*width+q];
        v = blockSum(v, smem);
        if (threadIdx.x == 0)
            out[q] = v;
    }
}

// Every column of a range of rows in one pass: per-thread sums of its rows, then one wavefront reduction per
// column and one barrier (a reduction per column, each with its barriers, made the final phase's sums slow).
constexpr int kMaxSumWidth = 12;

__global__ void __launch_bounds__(256) kRangeSums(const double* rows, int width, const int* starts,
                                                  const int* outIndex, int outStride, double* out) {
    __shared__ double smem[4][kMaxSumWidth];
    const int r = blockIdx.x;
    const int lo = starts[r], hi = starts[r+1];
    const int orow = outIndex ? outIndex[r] : r;
    double acc[kMaxSumWidth];
#pragma unroll
    for (int q = 0; q < kMaxSumWidth; q++)
        acc[q] = 0.0;
    for (int k = lo+threadIdx.x; k < hi; k += blockDim.x) {
        const double* row = rows+(size_t) k*width;
#pragma unroll
        for (int q = 0; q < kMaxSumWidth; q++)
            if (q < width)
                acc[q] += row[q];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kMaxSumWidth; q++) {
        const double v = groupSum<64>(acc[q]);
        if (lane == 0 && q < width)
            smem[w][q] = v;
    }
    __syncthreads();
    if ((int) threadIdx.x < width) {
        double v = 0.0;
        for (int k = 0; k < (int) (blockDim.x >> 6); k++)
            v += smem[k][threadIdx.x];
        out[(size_t) orow*outStride+threadIdx.x] = v;
    }
}

// The operands of a field evaluation of given charges q and dipoles d (caller's order): the sorted charges into
// the positions' w (FP64 and, when set, the pair type), the sorted dipoles as FP64 3-vectors and as the pair type
// with w = the radius parameter.
template<class T>
__global__ void kFieldOperands(int n, const int* order, const double* q, const double* d, const T* rgw,
                               double4* posD, typename Vec4<T>::type* posT, double* dS, typename Vec4<T>::type* dT) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const int i = order[s];
    posD[s].w = q[i];
    if (posT)
        posT[s].w = (T) q[i];
    typename Vec4<T>::type o;
    o.x = (T) d[3*i];
    o.y = (T) d[3*i+1];
    o.z = (T) d[3*i+2];
    o.w = rgw[s];
    dT[s] = o;
    for (int c = 0; c < 3; c++)
        dS[3*s+c] = d[3*i+c];
}

__global__ void kUnsort3(int n, const int* rank, const double* v, double* out) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= 3*n)
        return;
    const int i = k/3, c = k-3*i;
    out[k] = v[3*rank[i]+c];
}

}   // namespace

int cgBlocks(int n) { return std::max(1, std::min(kCgMaxBlocks, nblk(3LL*n, kCgBlock))); }

void launchGuess(hipStream_t st, int n, const int* order, const int* gidx, const double* hist, size_t slotStride,
                 GuessCoef gc, const double* alpha, const double* b, double* x) {
    kGuess<<<nblk(3LL*n, 256), 256, 0, st>>>(n, order, gidx, hist, slotStride, gc, alpha, b, x);
}

void launchCombineB(hipStream_t st, int n, const double* ed, const double* er, const double* perm, double cself,
                    double* b) {
    kCombineB<<<nblk(3LL*n, 256), 256, 0, st>>>(n, ed, er, perm, cself, b);
}

template<class T>
void launchToVec4(hipStream_t st, int n, const double* v, typename Vec4<T>::type* out, const T* rgw) {
    kToVec4<T><<<nblk(n, 256), 256, 0, st>>>(n, v, out, rgw);
}

template<class T>
void launchCgInit(hipStream_t st, int n, const double* alpha, const double* b, const double* x, const double* td,
                  const double* tr, double cself, double tol, int maxIter, double* r, double* z, double* p,
                  typename Vec4<T>::type* pT, double* partials, CgState* cg, int* done, Precond pc, const T* rgw) {
    const int nb = cgBlocks(n);
    kCgInit<T><<<nb, kCgBlock, 0, st>>>(n, alpha, b, x, td, tr, cself, r, z, p, pT, partials, pc, rgw);
    kCgInitCheck<<<1, kCgBlock, 0, st>>>(n, nb, tol, maxIter, partials, cg, done);
}

void launchCgAp(hipStream_t st, int n, int it, const double* alpha, const double* p, const double* td, const double* tr,
                double cself, double* ap, double* partials, const int* done) {
    kCgAp<<<cgBlocks(n), kCgBlock, 0, st>>>(n, alpha, p, td, tr, cself, ap, partials, done+it);
}

void launchCgUpdate(hipStream_t st, int n, int it, const double* alpha, const double* p, const double* ap, double* x,
                    double* r, double* z, double* partials, const int* done, Precond pc) {
    kCgUpdate<<<cgBlocks(n), kCgBlock, 0, st>>>(n, it, alpha, p, ap, x, r, z, partials, done+it, pc);
}

template<class T>
void launchPrecond(hipStream_t st, int n, int it, bool init, const double* alpha, const double* r, const double* ap,
                   const typename Vec4<T>::type* pos, BoxSel<T> box, const int* srCount, const int* srJ,
                   const typename Vec2<T>::type* srG, const int* over, double* z, double* p,
                   typename Vec4<T>::type* pT, double* partials, const int* done, const T* rgw) {
    const int nb = cgBlocks(n);
    auto go = [&](auto I, auto B) {
        kPrecond<T, I(), B()><<<nb, kCgBlock, 0, st>>>(n, it, alpha, r, ap, pos, box, srCount, srJ, srG, over, z, p,
                                                       pT, partials, done+it, rgw);
    };
    const bool b = box.table != nullptr;
    if (init)
        b ? go(std::true_type{}, std::true_type{}) : go(std::true_type{}, std::false_type{});
    else
        b ? go(std::false_type{}, std::true_type{}) : go(std::false_type{}, std::false_type{});
}

template<class T>
void launchCgDirection(hipStream_t st, int n, int it, double tol, int maxIter, const double* z, double* p,
                       typename Vec4<T>::type* pT, double* partials, CgState* cg, int* done, const T* rgw) {
    kCgDirection<T><<<cgBlocks(n), kCgBlock, 0, st>>>(n, it, tol, maxIter, z, p, pT, partials, cg, done, rgw);
}

template<class T>
void launchPost(hipStream_t st, int n, double peek, const double* alpha, const double* r, double* x, const double* perm,
                const int* order, const int* gidx, double* hist, double* d, const double* rgS,
                typename Vec4<T>::type* dT) {
    kPost<T><<<nblk(n, 256), 256, 0, st>>>(n, peek, alpha, r, x, perm, order, gidx, hist, d, rgS, dT);
}

void launchFinalCombine(hipStream_t st, int n, const double* d, const double* mu, const double* alpha,
                        const double* edir, const double* erec, double cself, double* etot, double* perAtom) {
    kFinalCombine<<<nblk(n, kCgBlock), kCgBlock, 0, st>>>(n, d, mu, alpha, edir, erec, cself, etot, perAtom);
}

void launchCovalentForces(hipStream_t st, int n, const int* rank, const double* pos, const int* covStart,
                          const int* covPartner, const double* covStrength, const int* revStart, const int* revEntry,
                          const int* covAtomOf, const double* etot, BoxSel<double> box, double* fcov,
                          double* perAtom) {
    if (box.table)
        kCovalent<true><<<nblk(n, kCgBlock), kCgBlock, 0, st>>>(n, rank, pos, covStart, covPartner, covStrength,
                                                                revStart, revEntry, covAtomOf, etot, box, fcov,
                                                                perAtom);
    else
        kCovalent<false><<<nblk(n, kCgBlock), kCgBlock, 0, st>>>(n, rank, pos, covStart, covPartner, covStrength,
                                                                 revStart, revEntry, covAtomOf, etot, box, fcov,
                                                                 perAtom);
}

void launchForcesOut(hipStream_t st, int n, const int* rank, const double* fpair, const double* frec,
                     const double* fcov, double ke, double* fOut, long long* fixedOut, int stride, double fixedScale,
                     const int* guard) {
    kForcesOut<<<nblk(n, 256), 256, 0, st>>>(n, rank, fpair, frec, fcov, ke, fOut, fixedOut, stride, fixedScale,
                                             guard);
}

void launchReducePartials(hipStream_t st, const double* partials, int nblk_, int width, double* out) {
    kReducePartials<<<1, 256, 0, st>>>(partials, nblk_, width, out);
}

void launchRangeSums(hipStream_t st, const double* rows, int width, const int* starts, int nr, const int* outIndex,
                     int outStride, double* out) {
    if (width < 1 || width > kMaxSumWidth)
        throw std::logic_error("pgm_hip: range sums of more than 12 columns");
    if (nr > 0)
        kRangeSums<<<nr, 256, 0, st>>>(rows, width, starts, outIndex, outStride, out);
}

template<class T>
void launchFieldOperands(hipStream_t st, int n, const int* order, const double* q, const double* d, const T* rgw,
                         double4* posD, typename Vec4<T>::type* posT, double* dS, typename Vec4<T>::type* dT) {
    kFieldOperands<T><<<nblk(n, 256), 256, 0, st>>>(n, order, q, d, rgw, posD, posT, dS, dT);
}

void launchUnsort3(hipStream_t st, int n, const int* rank, const double* v, double* out) {
    kUnsort3<<<nblk(3LL*n, 256), 256, 0, st>>>(n, rank, v, out);
}

#define PGM_INSTANTIATE_SOLVER(T)                                                                                     \
    template void launchFieldOperands<T>(hipStream_t, int, const int*, const double*, const double*, const T*,      \
                                         double4*, Vec4<T>::type*, double*, Vec4<T>::type*);                          \
    template void launchToVec4<T>(hipStream_t, int, const double*, Vec4<T>::type*, const T*);                       \
    template void launchCgInit<T>(hipStream_t, int, const double*, const double*, const double*, const double*,     \
                                  const double*, double, double, int, double*, double*, double*, Vec4<T>::type*,     \
                                  double*, CgState*, int*, Precond, const T*);                                       \
    template void launchPrecond<T>(hipStream_t, int, int, bool, const double*, const double*, const double*,         \
                                   const Vec4<T>::type*, BoxSel<T>, const int*, const int*, const Vec2<T>::type*,     \
                                   const int*, double*, double*, Vec4<T>::type*, double*, const int*, const T*);     \
    template void launchCgDirection<T>(hipStream_t, int, int, double, int, const double*, double*, Vec4<T>::type*,  \
                                       double*, CgState*, int*, const T*);                                           \
    template void launchPost<T>(hipStream_t, int, double, const double*, const double*, double*, const double*,     \
                                const int*, const int*, double*, double*, const double*, Vec4<T>::type*);
PGM_INSTANTIATE_SOLVER(float)
PGM_INSTANTIATE_SOLVER(double)

}   // namespace pgmhip
