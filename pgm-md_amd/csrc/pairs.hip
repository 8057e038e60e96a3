// Real-space pGM pair passes over the full neighbour lists (one group of L lanes per atom, no atomics):
//   setup  once per evaluation: compacts the candidates within the cutoff, stores (j, G1, G2) per entry and
//          returns the field of the charges and permanent dipoles (reference pairSetup / fieldPassDual,
//          core/cuda/pgm_cuda.cu:168-219, 1110-1185);
//   field  the dipole part of the CG matrix-vector product from the stored kernels (fieldPass,
//          core/cuda/pgm_cuda.cu:322-352);
//   final  energy, forces, field and strain at the total dipoles, the kernels recomputed (finalPass,
//          core/cuda/pgm_cuda.cu:354-438; core/pgm_core.cpp:422-440).
// T = float: FP32 pair arithmetic, FP64 accumulation per lane; T = double: FP64 throughout.
#include "kernels.hpp"

namespace pgmhip {

namespace {

constexpr int kPairLanes = 16;
constexpr int kPairBlock = 256;

template<class T, int L>
__global__ void __launch_bounds__(kPairBlock) kPairSetup(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                         const T* __restrict__ rg,
                                                         const typename Vec4<T>::type* __restrict__ dip,
                                                         const int* __restrict__ candStart,
                                                         const int* __restrict__ cand, int* __restrict__ nb,
                                                         int* __restrict__ pj, typename Vec2<T>::type* __restrict__ g12,
                                                         BoxT<T> box, T rc2, T beta, double* __restrict__ field) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    const int gid = blockIdx.x*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const int shift = (threadIdx.x%64)/L*L;
    const unsigned long long below = (lane == 0 ? 0ull : ((1ull << lane)-1ull));
    const unsigned long long gmask = (L == 64 ? ~0ull : ((1ull << L)-1ull));
    const T4 pi = pos[s];
    const T rgi = rg[s];
    const int start = candStart[s], cnt = candStart[s+1]-start;
    double ex = 0, ey = 0, ez = 0;
    int kept = 0;
    for (int base = 0; base < cnt; base += L) {
        const int k = base+lane;
        bool in = false;
        int j = 0;
        T G[3] = {T(0), T(0), T(0)};
        if (k < cnt) {
            j = cand[start+k];
            const T4 pjv = pos[j];
            T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
            minImage(x, y, z, box);
            const T r2 = x*x+y*y+z*z;
            in = r2 < rc2;
            if (in) {
                const T r = sqrt(r2);
                const T a = rsqrtT(rgi+rg[j]);
                pairKernels<T, 3>(a, beta, r, G);
                const T4 d = dip[j];
                const T djx = d.x*x+d.y*y+d.z*z;
                const T c1 = pjv.w*G[1], c2 = djx*G[2];
                ex += (double) ((c1+c2)*x-d.x*G[1]);
                ey += (double) ((c1+c2)*y-d.y*G[1]);
                ez += (double) ((c1+c2)*z-d.z*G[1]);
            }
        }
        const unsigned long long m = (__ballot(in) >> shift) & gmask;
        if (in) {
            const int idx = start+kept+__popcll(m & below);
            pj[idx] = j;
            T2 g;
            g.x = G[1];
            g.y = G[2];
            g12[idx] = g;
        }
        kept += __popcll(m);
    }
    ex = groupSum<L>(ex);
    ey = groupSum<L>(ey);
    ez = groupSum<L>(ez);
    if (lane == 0) {
        field[3*s] = ex;
        field[3*s+1] = ey;
        field[3*s+2] = ez;
        nb[s] = kept;
    }
}

template<class T, int L>
__global__ void __launch_bounds__(kPairBlock) kPairField(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                         const typename Vec4<T>::type* __restrict__ dip,
                                                         const int* __restrict__ candStart,
                                                         const int* __restrict__ nb, const int* __restrict__ pj,
                                                         const typename Vec2<T>::type* __restrict__ g12, BoxT<T> box,
                                                         double* __restrict__ field, const int* skip) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    if (skip && *skip)
        return;
    const int gid = blockIdx.x*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const T4 pi = pos[s];
    const int start = candStart[s], cnt = nb[s];
    double ex = 0, ey = 0, ez = 0;
    for (int k = lane; k < cnt; k += L) {
        const int j = pj[start+k];
        const T2 g = g12[start+k];
        const T4 pjv = pos[j];
        T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
        minImage(x, y, z, box);
        const T4 d = dip[j];
        const T c2 = (d.x*x+d.y*y+d.z*z)*g.y;
        ex += (double) (c2*x-d.x*g.x);
        ey += (double) (c2*y-d.y*g.x);
        ez += (double) (c2*z-d.z*g.x);
    }
    ex = groupSum<L>(ex);
    ey = groupSum<L>(ey);
    ez = groupSum<L>(ez);
    if (lane == 0) {
        field[3*s] = ex;
        field[3*s+1] = ey;
        field[3*s+2] = ez;
    }
}

template<class T, int L>
__global__ void __launch_bounds__(kPairBlock) kPairFinal(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                         const T* __restrict__ rg,
                                                         const typename Vec4<T>::type* __restrict__ dip,
                                                         const int* __restrict__ candStart,
                                                         const int* __restrict__ nb, const int* __restrict__ pj,
                                                         BoxT<T> box, T beta, double* __restrict__ force,
                                                         double* __restrict__ field, double* __restrict__ partials) {
    typedef typename Vec4<T>::type T4;
    __shared__ double red[kPairBlock/64][10];
    const int gid = blockIdx.x*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // energy, strain (row-major)
    if (s < n) {
        const T4 pi = pos[s];
        const T4 di = dip[s];
        const T qi = pi.w, rgi = rg[s];
        const int start = candStart[s], cnt = nb[s];
        double fx = 0, fy = 0, fz = 0, ex = 0, ey = 0, ez = 0;
        for (int k = lane; k < cnt; k += L) {
            const int j = pj[start+k];
            const T4 pjv = pos[j];
            T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
            minImage(x, y, z, box);
            const T r = sqrt(x*x+y*y+z*z);
            const T a = rsqrtT(rgi+rg[j]);
            T G[4];
            pairKernels<T, 4>(a, beta, r, G);
            const T4 dj = dip[j];
            const T qj = pjv.w;
            const T dix = di.x*x+di.y*y+di.z*z, djx = dj.x*x+dj.y*y+dj.z*z, didj = di.x*dj.x+di.y*dj.y+di.z*dj.z;
            const T qd = qi*djx-qj*dix;
            acc[0] += (double) (qi*qj*G[0]+(qd+didj)*G[1]-dix*djx*G[2]);
            const T sc = -qi*qj*G[1]-(qd+didj)*G[2]+dix*djx*G[3];
            const T gx = sc*x+(qi*dj.x-qj*di.x)*G[1]-(djx*di.x+dix*dj.x)*G[2];
            const T gy = sc*y+(qi*dj.y-qj*di.y)*G[1]-(djx*di.y+dix*dj.y)*G[2];
            const T gz = sc*z+(qi*dj.z-qj*di.z)*G[1]-(djx*di.z+dix*dj.z)*G[2];
            fx -= (double) gx;
            fy -= (double) gy;
            fz -= (double) gz;
            acc[1] += (double) (gx*x);
            acc[2] += (double) (gx*y);
            acc[3] += (double) (gx*z);
            acc[4] += (double) (gy*x);
            acc[5] += (double) (gy*y);
            acc[6] += (double) (gy*z);
            acc[7] += (double) (gz*x);
            acc[8] += (double) (gz*y);
            acc[9] += (double) (gz*z);
            const T c = qj*G[1]+djx*G[2];
            ex += (double) (c*x-dj.x*G[1]);
            ey += (double) (c*y-dj.y*G[1]);
            ez += (double) (c*z-dj.z*G[1]);
        }
        fx = groupSum<L>(fx);
        fy = groupSum<L>(fy);
        fz = groupSum<L>(fz);
        ex = groupSum<L>(ex);
        ey = groupSum<L>(ey);
        ez = groupSum<L>(ez);
        if (lane == 0) {
            force[3*s] = fx;
            force[3*s+1] = fy;
            force[3*s+2] = fz;
            field[3*s] = ex;
            field[3*s+1] = ey;
            field[3*s+2] = ez;
        }
    }
    // each pair is visited from both atoms: half of the sums
#pragma unroll
    for (int q = 0; q < 10; q++) {
        const double v = groupSum<64>(acc[q]);
        if ((threadIdx.x & 63) == 0)
            red[threadIdx.x >> 6][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        double v = 0;
        for (int w = 0; w < kPairBlock/64; w++)
            v += red[w][threadIdx.x];
        partials[(size_t) blockIdx.x*10+threadIdx.x] = 0.5*v;
    }
}

inline int groupsBlocks(int n, int L) { return (int) (((long long) n*L+kPairBlock-1)/kPairBlock); }

}   // namespace

template<class T>
void launchPairSetup(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg,
                     const typename Vec4<T>::type* dip, const int* candStart, const int* cand, int* nb, int* pj,
                     typename Vec2<T>::type* g12, BoxT<T> box, T rc2, T beta, double* field, const int*) {
    kPairSetup<T, kPairLanes><<<groupsBlocks(n, kPairLanes), kPairBlock, 0, st>>>(n, pos, rg, dip, candStart, cand, nb,
                                                                                 pj, g12, box, rc2, beta, field);
}

template<class T>
void launchPairField(hipStream_t st, int n, const typename Vec4<T>::type* pos, const typename Vec4<T>::type* dip,
                     const int* candStart, const int* nb, const int* pj, const typename Vec2<T>::type* g12,
                     BoxT<T> box, double* field, const int* skip) {
    kPairField<T, kPairLanes><<<groupsBlocks(n, kPairLanes), kPairBlock, 0, st>>>(n, pos, dip, candStart, nb, pj, g12,
                                                                                 box, field, skip);
}

template<class T>
int launchPairFinal(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg,
                    const typename Vec4<T>::type* dip, const int* candStart, const int* nb, const int* pj, BoxT<T> box,
                    T beta, double* force, double* field, double* partials) {
    const int blocks = groupsBlocks(n, kPairLanes);
    kPairFinal<T, kPairLanes><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, dip, candStart, nb, pj, box, beta, force,
                                                            field, partials);
    return blocks;
}

#define PGM_INSTANTIATE_PAIRS(T)                                                                                      \
    template void launchPairSetup<T>(hipStream_t, int, const Vec4<T>::type*, const T*, const Vec4<T>::type*,         \
                                     const int*, const int*, int*, int*, Vec2<T>::type*, BoxT<T>, T, T, double*,      \
                                     const int*);                                                                     \
    template void launchPairField<T>(hipStream_t, int, const Vec4<T>::type*, const Vec4<T>::type*, const int*,       \
                                     const int*, const int*, const Vec2<T>::type*, BoxT<T>, double*, const int*);     \
    template int launchPairFinal<T>(hipStream_t, int, const Vec4<T>::type*, const T*, const Vec4<T>::type*,          \
                                    const int*, const int*, const int*, BoxT<T>, T, double*, double*, double*);
PGM_INSTANTIATE_PAIRS(float)
PGM_INSTANTIATE_PAIRS(double)

}   // namespace pgmhip
