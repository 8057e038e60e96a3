// Real-space pGM pair passes over the full neighbour lists, one group of L = 16 lanes per atom, no atomics:
//   setup  once per evaluation: compacts the candidates within the cutoff (a cheap pass), then evaluates G0..G3
//          over the compacted pairs with every lane busy, stores (j, G1, G2) and (G0, G3) per pair, and returns
//          the field of the charges and permanent dipoles and, dual, the first CG product T x0 (reference
//          pairSetup / pairSetupDual, core/cuda/pgm_cuda.cu:168-219, 1110-1185);
//   field  the dipole part of a CG matrix-vector product from the stored G1, G2 (fieldPass,
//          core/cuda/pgm_cuda.cu:322-352);
//   final  energy, forces, field and strain at the total dipoles from the stored G0..G3 (finalPass,
//          core/cuda/pgm_cuda.cu:354-438; core/pgm_core.cpp:422-440).
// The field and final passes stream the list: every lane first issues the loads of U entries (index and
// kernels, then the partners' positions and dipoles), then computes, so U independent gathers are in flight.
// T = float: FP32 pair arithmetic with FP64 accumulation per lane; T = double: FP64 throughout.
#include "kernels.hpp"

#include <stdexcept>
#include <type_traits>

namespace pgmhip {

namespace {

constexpr int kPairLanes = 16;
constexpr int kPairBlock = 256;

// FP64 stores all four kernels per pair; FP32 stores G1, G2 only and rebuilds G0, G3 where they are needed
// (g03FromG12): the setup pass writes 8 bytes per pair instead of 16, the final pass reads 12 instead of 20.
template<class T>
constexpr bool kStoreG03 = sizeof(T) == sizeof(double);

// The separate setup pass writes the stored kernels (FP64: 32 bytes per pair, 2.8 GB per evaluation at 96k atoms)
// with non-temporal stores: no pass reads them again soon enough to find them in a cache, and the data written
// normally, the compacted list that every field pass reads again, keeps its place there.
template<class V>
__device__ __forceinline__ void streamStore(V* p, const V& v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// Compaction: every atom's candidates within rc, in place in its segment (pj), and the count (nb).
template<class T, int L, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kCompact(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                       const int* __restrict__ candStart,
                                                       const int* __restrict__ cand, int* __restrict__ nb,
                                                       int* __restrict__ pj, BoxSel<T> bs, T rc2) {
    typedef typename Vec4<T>::type T4;
    const int gid = xcdBlock()*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const BoxT<T> box = selectBox<BATCH>(bs, s);
    const int shift = (threadIdx.x%64)/L*L;
    const unsigned long long below = (lane == 0 ? 0ull : ((1ull << lane)-1ull));
    const unsigned long long gmask = (L == 64 ? ~0ull : ((1ull << (L & 63))-1ull));
    const T4 pi = pos[s];
    const int start = candStart[s], cnt = candStart[s+1]-start;
    int kept = 0;
    for (int base = 0; base < cnt; base += L) {
        const int k = base+lane;
        bool in = false;
        int j = 0;
        if (k < cnt) {
            j = cand[start+k];   // the raw entry: a caller's list carries the image in it
            const T4 pjv = pos[entryAtom(j, bs.ext)];
            T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
            pairImage(x, y, z, box, j, bs.ext);
            in = x*x+y*y+z*z < rc2;
        }
        const unsigned long long m = (__ballot(in) >> shift) & gmask;
        if (in)
            pj[start+kept+__popcll(m & below)] = j;
        kept += __popcll(m);
    }
    if (lane == 0)
        nb[s] = kept;
}

// The kernels G0..G3 of one compacted pair, the field of (q, dip) at i and, DUAL, of dual at i.
template<class T, bool DUAL>
__device__ __forceinline__ void setupTerm(const typename Vec4<T>::type& pi, T rgi, const typename Vec4<T>::type& pjv,
                                          T rgj, const typename Vec4<T>::type& d, const typename Vec4<T>::type& u,
                                          const BoxT<T>& box, int v, int ext, T beta, typename Vec2<T>::type& ga,
                                          typename Vec2<T>::type& gb, double* e, double* f) {
    T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
    pairImage(x, y, z, box, v, ext);
    T G[4];
    if constexpr (sizeof(T) == sizeof(float))
        pairKernelsF<4>(rgi+rgj, beta, x*x+y*y+z*z, G);
    else
        pairKernels<T, 4>(rsqrtT(rgi+rgj), beta, sqrt(x*x+y*y+z*z), G);
    const T c = pjv.w*G[1]+(d.x*x+d.y*y+d.z*z)*G[2];
    e[0] += (double) (c*x-d.x*G[1]);
    e[1] += (double) (c*y-d.y*G[1]);
    e[2] += (double) (c*z-d.z*G[1]);
    if (DUAL) {
        const T c2 = (u.x*x+u.y*y+u.z*z)*G[2];
        f[0] += (double) (c2*x-u.x*G[1]);
        f[1] += (double) (c2*y-u.y*G[1]);
        f[2] += (double) (c2*z-u.z*G[1]);
    }
    ga.x = G[1];
    ga.y = G[2];
    gb.x = G[0];
    gb.y = G[3];
}

// setupTerm from the displacement x = r_i - r_j (the pair's image) and the partner's charge.
template<class T, bool DUAL>
__device__ __forceinline__ void setupTermX(T x, T y, T z, T qj, T rgSum, const typename Vec4<T>::type& d,
                                           const typename Vec4<T>::type& u, T beta, typename Vec2<T>::type& ga,
                                           typename Vec2<T>::type& gb, double* e, double* f) {
    T G[4];
    if constexpr (sizeof(T) == sizeof(float))
        pairKernelsF<4>(rgSum, beta, x*x+y*y+z*z, G);
    else
        pairKernels<T, 4>(rsqrtT(rgSum), beta, sqrt(x*x+y*y+z*z), G);
    const T c = qj*G[1]+(d.x*x+d.y*y+d.z*z)*G[2];
    e[0] += (double) (c*x-d.x*G[1]);
    e[1] += (double) (c*y-d.y*G[1]);
    e[2] += (double) (c*z-d.z*G[1]);
    if (DUAL) {
        const T c2 = (u.x*x+u.y*y+u.z*z)*G[2];
        f[0] += (double) (c2*x-u.x*G[1]);
        f[1] += (double) (c2*y-u.y*G[1]);
        f[2] += (double) (c2*z-u.z*G[1]);
    }
    ga.x = G[1];
    ga.y = G[2];
    gb.x = G[0];
    gb.y = G[3];
}

// LDS written by some lanes of a wavefront and read by others: no reordering across this point, and the
// wavefront's LDS operations complete in order.
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Compaction and setup in one pass.  Each group of L lanes tests L candidates per step (one per lane) and
// ballot-compacts those within rc, with their minimum-image displacement and charge, into its LDS buffer of 2L
// entries; whenever a group of the wavefront has L entries buffered, every group evaluates up to L of its own
// (kernels, fields) and writes (j, G1, G2) as contiguous entries of its atom's segment.  No compacted index list
// goes through HBM, the partner positions are gathered once, and the transcendental part runs with every lane
// busy.
//
// Close pairs (a r < 1, r^2 < rg_i + rg_j: bonded neighbours and close contacts, about 2 per atom in water) take
// the series form of the kernels, which every lane of a wavefront steps through when one of its lanes holds such
// a pair.  Candidates with r^2 < rg_i + rgMax (rgMax the largest rg, so a superset of the close pairs) are held
// apart in a stash of L entries per group and evaluated together after the others, as the last entries of the
// segment: the evaluations here and the passes over the list meet the series form in the last row of a segment
// instead of in a few rows spread over it.  A stash that is full leaves further ones in the stream (the order
// is a matter of speed only; every entry is evaluated by the same code whatever its place).
template<class T, int L, bool DUAL, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kCompactSetup(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                            const T* __restrict__ rg,
                                                            const typename Vec4<T>::type* __restrict__ dip,
                                                            const typename Vec4<T>::type* __restrict__ dual,
                                                            const int* __restrict__ candStart,
                                                            const int* __restrict__ cand, int* __restrict__ nb,
                                                            int* __restrict__ pj,
                                                            typename Vec2<T>::type* __restrict__ g12,
                                                            typename Vec2<T>::type* __restrict__ g03, BoxSel<T> bs,
                                                            T rc2, T beta, T rgMax, double* __restrict__ field,
                                                            double* __restrict__ fieldDual) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    __shared__ int bufJ[2*kPairBlock];
    __shared__ T4 bufX[2*kPairBlock];
    __shared__ int closeJ[kPairBlock];
    __shared__ T4 closeX[kPairBlock];
    const int gid = xcdBlock()*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const int g0 = 2*(threadIdx.x-lane);   // this group's 2L buffer slots
    const int c0 = threadIdx.x-lane;       // and its L stash slots
    const BoxT<T> box = selectBox<BATCH>(bs, s);
    const int shift = (threadIdx.x%64)/L*L;
    const unsigned long long below = (lane == 0 ? 0ull : ((1ull << lane)-1ull));
    const unsigned long long gmask = (L == 64 ? ~0ull : ((1ull << (L & 63))-1ull));
    const T4 pi = pos[s];
    const T rgi = rg[s];
    const T closeR2 = rgi+rgMax;
    const int start = candStart[s], cnt = candStart[s+1]-start;
    double e[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    int kept = 0, buffered = 0, stashed = 0;
    auto evaluate = [&](int v, const T4& xq, int out) {   // v the raw entry (stored as is); j its atom
        const int j = entryAtom(v, bs.ext);
        const T4 d = dip[j];
        const T4 u = DUAL ? dual[j] : d;
        T2 ga, gb;
        setupTermX<T, DUAL>(xq.x, xq.y, xq.z, xq.w, rgi+d.w, d, u, beta, ga, gb, e, f);
        pj[out] = v;
        g12[out] = ga;
        if (kStoreG03<T>)
            g03[out] = gb;
    };
    // software pipeline: the candidate index two chunks ahead and the position one chunk ahead are in flight
    // while a chunk is tested (and a flush evaluated)
    int jNext = (lane < cnt) ? cand[start+lane] : 0;
    int jAfter = (L+lane < cnt) ? cand[start+L+lane] : 0;
    T4 pNext;
    if (lane < cnt)
        pNext = pos[entryAtom(jNext, bs.ext)];
    for (int first = 0; first < cnt; first += L) {
        const int k = first+lane;
        bool in = false;
        const int j = jNext;
        const T4 pjv = pNext;
        jNext = jAfter;
        if (k+L < cnt)
            pNext = pos[entryAtom(jNext, bs.ext)];
        if (k+2*L < cnt)
            jAfter = cand[start+k+2*L];
        T4 xq;
        bool close = false;
        if (k < cnt) {
            T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
            pairImage(x, y, z, box, j, bs.ext);
            const T r2 = x*x+y*y+z*z;
            in = r2 < rc2;
            close = in && r2 < closeR2;
            xq.x = x;
            xq.y = y;
            xq.z = z;
            xq.w = pjv.w;
        }
        const unsigned long long mc = (__ballot(close) >> shift) & gmask;
        if (mc != 0ull) {   // group-uniform
            const int rank = __popcll(mc & below);
            if (close && stashed+rank < L) {
                closeJ[c0+stashed+rank] = j;
                closeX[c0+stashed+rank] = xq;
                in = false;
            }
            stashed = min(stashed+__popcll(mc), L);
        }
        const unsigned long long m = (__ballot(in) >> shift) & gmask;
        if (in) {
            const int slot = g0+buffered+__popcll(m & below);
            bufJ[slot] = j;
            bufX[slot] = xq;
        }
        buffered += __popcll(m);
        waveSync();
        // when any group of the wavefront has L entries, every group flushes up to L: the wavefront evaluates
        // once for all its groups instead of once per group
        if (__ballot(lane == 0 && buffered >= L) != 0ull) {
            const int nflush = min(buffered, L);
            if (lane < nflush)
                evaluate(bufJ[g0+lane], bufX[g0+lane], start+kept+lane);
            kept += nflush;
            const int rest = buffered-nflush;
            int tj = 0;
            T4 tx;
            if (lane < rest) {
                tj = bufJ[g0+nflush+lane];
                tx = bufX[g0+nflush+lane];
            }
            waveSync();
            if (lane < rest) {
                bufJ[g0+lane] = tj;
                bufX[g0+lane] = tx;
            }
            waveSync();
            buffered = rest;
        }
    }
    if (lane < buffered)
        evaluate(bufJ[g0+lane], bufX[g0+lane], start+kept+lane);
    kept += buffered;
    if (lane < stashed)
        evaluate(closeJ[c0+lane], closeX[c0+lane], start+kept+lane);
    kept += stashed;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        e[c] = groupSum<L>(e[c]);
        if (DUAL)
            f[c] = groupSum<L>(f[c]);
    }
    if (lane == 0) {
        nb[s] = kept;
        field[3*s] = e[0];
        field[3*s+1] = e[1];
        field[3*s+2] = e[2];
        if (DUAL) {
            fieldDual[3*s] = f[0];
            fieldDual[3*s+1] = f[1];
            fieldDual[3*s+2] = f[2];
        }
    }
}

// Kernels and fields over the compacted pairs: every lane busy, U entries in flight per lane.
template<class T, int L, int U, bool DUAL, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kPairSetup(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                         const T* __restrict__ rg,
                                                         const typename Vec4<T>::type* __restrict__ dip,
                                                         const typename Vec4<T>::type* __restrict__ dual,
                                                         const int* __restrict__ candStart,
                                                         const int* __restrict__ nb, const int* __restrict__ pj,
                                                         typename Vec2<T>::type* __restrict__ g12,
                                                         typename Vec2<T>::type* __restrict__ g03, BoxSel<T> bs, T beta,
                                                         double* __restrict__ field, double* __restrict__ fieldDual) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    const int gid = xcdBlock()*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const BoxT<T> box = selectBox<BATCH>(bs, s);
    const T4 pi = pos[s];
    const T rgi = rg[s];
    const int start = candStart[s], cnt = nb[s];
    double e[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    int k = lane;
    for (; k+(U-1)*L < cnt; k += U*L) {
        int v[U], j[U];
        T4 pjv[U], d[U], u[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            v[q] = pj[start+k+q*L];
            j[q] = entryAtom(v[q], bs.ext);
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            pjv[q] = pos[j[q]];
            d[q] = dip[j[q]];
            if (DUAL)
                u[q] = dual[j[q]];
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            T2 ga, gb;
            setupTerm<T, DUAL>(pi, rgi, pjv[q], d[q].w, d[q], u[q], box, v[q], bs.ext, beta, ga, gb, e, f);
            streamStore(g12+start+k+q*L, ga);
            if (kStoreG03<T>)
                streamStore(g03+start+k+q*L, gb);
        }
    }
    for (; k < cnt; k += L) {
        const int v = pj[start+k], j = entryAtom(v, bs.ext);
        T2 ga, gb;
        const T4 u = DUAL ? dual[j] : pos[j];
        const T4 dj = dip[j];
        setupTerm<T, DUAL>(pi, rgi, pos[j], dj.w, dj, u, box, v, bs.ext, beta, ga, gb, e, f);
        streamStore(g12+start+k, ga);
        if (kStoreG03<T>)
            streamStore(g03+start+k, gb);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        e[c] = groupSum<L>(e[c]);
        if (DUAL)
            f[c] = groupSum<L>(f[c]);
    }
    if (lane == 0) {
        field[3*s] = e[0];
        field[3*s+1] = e[1];
        field[3*s+2] = e[2];
        if (DUAL) {
            fieldDual[3*s] = f[0];
            fieldDual[3*s+1] = f[1];
            fieldDual[3*s+2] = f[2];
        }
    }
}

template<class T>
__device__ __forceinline__ void fieldTerm(const typename Vec4<T>::type& pi, const typename Vec4<T>::type& pj,
                                          const typename Vec4<T>::type& d, const typename Vec2<T>::type& g,
                                          const BoxT<T>& box, int v, int ext, double& ex, double& ey,
                                          double& ez) {
    T x = pi.x-pj.x, y = pi.y-pj.y, z = pi.z-pj.z;
    pairImage(x, y, z, box, v, ext);
    const T c2 = (d.x*x+d.y*y+d.z*z)*g.y;
    ex += (double) (c2*x-d.x*g.x);
    ey += (double) (c2*y-d.y*g.x);
    ez += (double) (c2*z-d.z*g.x);
}

template<class T, int L, int U, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kPairField(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                         const typename Vec4<T>::type* __restrict__ dip,
                                                         const int* __restrict__ candStart,
                                                         const int* __restrict__ nb, const int* __restrict__ pj,
                                                         const typename Vec2<T>::type* __restrict__ g12, BoxSel<T> bs,
                                                         double* __restrict__ field, const int* skip,
                                                         bool reverse) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    if (skip && *skip)
        return;
    const int gid = xcdBlock(reverse)*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const BoxT<T> box = selectBox<BATCH>(bs, s);
    const T4 pi = pos[s];
    const int start = candStart[s], cnt = nb[s];
    double ex = 0, ey = 0, ez = 0;
    int k = lane;
    for (; k+(U-1)*L < cnt; k += U*L) {
        int v[U];
        T2 g[U];
        T4 pjv[U], d[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            v[u] = pj[start+k+u*L];
            g[u] = g12[start+k+u*L];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = entryAtom(v[u], bs.ext);
            pjv[u] = pos[j];
            d[u] = dip[j];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            fieldTerm<T>(pi, pjv[u], d[u], g[u], box, v[u], bs.ext, ex, ey, ez);
    }
    for (; k < cnt; k += L) {
        const int v = pj[start+k], j = entryAtom(v, bs.ext);
        fieldTerm<T>(pi, pos[j], dip[j], g12[start+k], box, v, bs.ext, ex, ey, ez);
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

// A buffer resource over n float4 records (the gfx9 raw-buffer data word): a gather then addresses a record with
// one 32-bit offset, the same for the position and the dipole of a partner, instead of two 64-bit addresses.  An
// offset past the records reads zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t float4Records(const float4* p, int n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(p), (short) 0, n*(int) sizeof(float4), 0x00020000);
}
__device__ __forceinline__ float4 loadRecord(__amdgpu_buffer_rsrc_t r, int offset) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, offset, 0, 0));
}

// The FP32 field pass with the kernels evaluated on the fly (no stored kernels read): per entry the partner's
// index, position and dipole (w = its radius parameter rg_j; rg_i rides in the atom's own dipole w), G1 and G2
// from r^2 and rg_i + rg_j (pairKernelsF, as the setup pass evaluates them).
template<int L, int U, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kPairFieldOtf(int n, const float4* __restrict__ pos,
                                                            const float4* __restrict__ dip,
                                                            const int* __restrict__ candStart,
                                                            const int* __restrict__ nb, const int* __restrict__ pj,
                                                            BoxSel<float> bs, float beta, double* __restrict__ field,
                                                            const int* skip, bool reverse) {
    if (skip && *skip)
        return;
    const int gid = xcdBlock(reverse)*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const BoxT<float> box = selectBox<BATCH>(bs, s);
    const float4 pi = pos[s];
    const float rgi = dip[s].w;
    const __amdgpu_buffer_rsrc_t posR = float4Records(pos, n), dipR = float4Records(dip, n);
    const int start = candStart[s], cnt = nb[s];
    // the U entries of a step summed in FP32 (the matrix-vector product is FP32), the steps in FP64
    double ex = 0, ey = 0, ez = 0;
    auto term = [&](const float4& pjv, const float4& d, int v, float& fx, float& fy, float& fz) {
        float x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
        pairImage(x, y, z, box, v, bs.ext);
        float G[3];
        pairKernelsF<3>(rgi+d.w, beta, x*x+y*y+z*z, G);
        const float c2 = (d.x*x+d.y*y+d.z*z)*G[2];
        fx = fmaf(c2, x, fmaf(-d.x, G[1], fx));
        fy = fmaf(c2, y, fmaf(-d.y, G[1], fy));
        fz = fmaf(c2, z, fmaf(-d.z, G[1], fz));
    };
    int k = lane;
    for (; k+(U-1)*L < cnt; k += U*L) {
        int v[U];
        float4 pjv[U], d[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            v[u] = pj[start+k+u*L];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int o = entryAtom(v[u], bs.ext)*(int) sizeof(float4);
            pjv[u] = loadRecord(posR, o);
            d[u] = loadRecord(dipR, o);
        }
        float fx = -0.0f, fy = -0.0f, fz = -0.0f;   // -0: adding it is exact for every x, so it folds away
#pragma unroll
        for (int u = 0; u < U; u++)
            term(pjv[u], d[u], v[u], fx, fy, fz);
        ex += (double) fx;
        ey += (double) fy;
        ez += (double) fz;
    }
    for (; k < cnt; k += L) {
        const int v = pj[start+k], o = entryAtom(v, bs.ext)*(int) sizeof(float4);
        float fx = -0.0f, fy = -0.0f, fz = -0.0f;   // -0: adding it is exact for every x, so it folds away
        term(loadRecord(posR, o), loadRecord(dipR, o), v, fx, fy, fz);
        ex += (double) fx;
        ey += (double) fy;
        ez += (double) fz;
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

// The GVDW pair term of one visit of pair (i, j) (combination rule: A = a_i a_j, C6 = c_i c_j, B = (b_i + b_j)/2
// from gv = (a, c, b), b_ij = (rg_i + rg_j)^-1/2): energy into eg, force on i and, STRAIN, the strain into f and
// acc in units of ke (invKe = 1/ke), as the electrostatic terms (core/cuda/pgm_cuda.cu:398-414).  sign = -1
// removes exactly what sign = +1 added: the skipped pairs.
template<class T, bool STRAIN, class F>
__device__ __forceinline__ void gvdwTerm(T x, T y, T z, T r2, const typename Vec4<T>::type& gvi,
                                         const typename Vec4<T>::type& gvj, T rgSum, T rvdw2, bool slater,
                                         T invKe, double sign, double& eg, double* acc, F* f) {
    const T A = gvi.x*gvj.x, C6 = gvi.y*gvj.y;
    if (!(r2 < rvdw2) || (A == T(0) && C6 == T(0)))
        return;
    T dEdR;
    const T u = gvdwPair<T>(sqrt(r2), rsqrtT(rgSum), A, C6, T(0.5)*(gvi.z+gvj.z), slater, dEdR);
    const T fs = dEdR*invKe;
    eg += sign*(double) u;
    f[0] += (F) (sign*(double) (fs*x));
    f[1] += (F) (sign*(double) (fs*y));
    f[2] += (F) (sign*(double) (fs*z));
    if (STRAIN) {
        const T xs[3] = {x, y, z};
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                acc[1+3*a+c] -= sign*(double) (fs*xs[a]*xs[c]);
    }
}

// One pair of the final pass at fixed total dipoles: energy, dE/dx, field at i and, STRAIN, the strain
// (core/pgm_core.cpp:422-440); GVDW adds the pair's GVDW term.
template<class T, bool STRAIN, bool GVDW, class A>
__device__ __forceinline__ void finalTerm(const typename Vec4<T>::type& pi, const typename Vec4<T>::type& di,
                                          const typename Vec4<T>::type& pjv, const typename Vec4<T>::type& dj,
                                          const typename Vec2<T>::type& ga, const typename Vec2<T>::type& gb,
                                          T rgSum, T beta, const BoxT<T>& box, int v, int ext, double* acc, A* f,
                                          A* e, A& u0, const typename Vec4<T>::type& gvi,
                                          const typename Vec4<T>::type& gvj, const GvdwArgs<T>& gva, double& eg) {
    T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
    pairImage(x, y, z, box, v, ext);
    if (GVDW)
        gvdwTerm<T, STRAIN>(x, y, z, x*x+y*y+z*z, gvi, gvj, rgSum, gva.rvdw2, gva.slater != 0, gva.invKe, 1.0, eg,
                            acc, f);
    const T G1 = ga.x, G2 = ga.y;
    T G0, G3;
    if constexpr (kStoreG03<T>) {
        G0 = gb.x;
        G3 = gb.y;
        (void) beta;
    }
    else
        g03FromG12(G1, G2, x*x+y*y+z*z, rgSum, beta, G0, G3);
    const T qi = pi.w, qj = pjv.w;
    const T dix = di.x*x+di.y*y+di.z*z, djx = dj.x*x+dj.y*y+dj.z*z, didj = di.x*dj.x+di.y*dj.y+di.z*dj.z;
    const T qd = qi*djx-qj*dix;
    u0 += (A) (qi*qj*G0+(qd+didj)*G1-dix*djx*G2);
    const T sc = -qi*qj*G1-(qd+didj)*G2+dix*djx*G3;
    const T gx = sc*x+(qi*dj.x-qj*di.x)*G1-(djx*di.x+dix*dj.x)*G2;
    const T gy = sc*y+(qi*dj.y-qj*di.y)*G1-(djx*di.y+dix*dj.y)*G2;
    const T gz = sc*z+(qi*dj.z-qj*di.z)*G1-(djx*di.z+dix*dj.z)*G2;
    f[0] -= (A) gx;
    f[1] -= (A) gy;
    f[2] -= (A) gz;
    if (STRAIN) {
        acc[1] += (double) (gx*x);
        acc[2] += (double) (gx*y);
        acc[3] += (double) (gx*z);
        acc[4] += (double) (gy*x);
        acc[5] += (double) (gy*y);
        acc[6] += (double) (gy*z);
        acc[7] += (double) (gz*x);
        acc[8] += (double) (gz*y);
        acc[9] += (double) (gz*z);
    }
    const T c = qj*G1+djx*G2;
    e[0] += (A) (c*x-dj.x*G1);
    e[1] += (A) (c*y-dj.y*G1);
    e[2] += (A) (c*z-dj.z*G1);
}

// perAtom: [n][10] (energy, strain row-major) with STRAIN, else [n] (energy), each half of the atom's pair sums;
// GVDW: the GVDW energy per atom into gva.energy.
template<class T, int L, int U, bool STRAIN, bool GVDW, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kPairFinal(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                         const typename Vec4<T>::type* __restrict__ dip,
                                                         const T* __restrict__ rg, T beta,
                                                         const int* __restrict__ candStart,
                                                         const int* __restrict__ nb, const int* __restrict__ pj,
                                                         const typename Vec2<T>::type* __restrict__ g12,
                                                         const typename Vec2<T>::type* __restrict__ g03, BoxSel<T> bs,
                                                         double* __restrict__ force, double* __restrict__ field,
                                                         double* __restrict__ perAtom, GvdwArgs<T> gva,
                                                         bool reverse) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    const int gid = xcdBlock(reverse)*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    constexpr int NACC = STRAIN ? 10 : 1;
    double eg = 0;
    double acc[NACC];   // energy, strain (row-major)
#pragma unroll
    for (int q = 0; q < NACC; q++)
        acc[q] = 0;
    if (s < n) {
        const BoxT<T> box = selectBox<BATCH>(bs, s);
        const T4 pi = pos[s];
        const T4 di = dip[s];
        const T rgi = (kStoreG03<T> && !GVDW) ? T(0) : rg[s];
        const T4 gvi = GVDW ? gva.gv[s] : di;
        const int start = candStart[s], cnt = nb[s];
        // FP32: the partner gathers through buffer resources (as in kPairFieldOtf); unused for FP64
        const __amdgpu_buffer_rsrc_t posR = float4Records(reinterpret_cast<const float4*>(pos), n);
        const __amdgpu_buffer_rsrc_t dipR = float4Records(reinterpret_cast<const float4*>(dip), n);
        auto gather = [&](int jj, T4& p, T4& d) {
            if constexpr (sizeof(T) == sizeof(float)) {
                const int o = jj*(int) sizeof(float4);
                p = loadRecord(posR, o);
                d = loadRecord(dipR, o);
            }
            else {
                p = pos[jj];
                d = dip[jj];
            }
        };
        double f[3] = {0, 0, 0}, e[3] = {0, 0, 0};
        int k = lane;
        for (; k+(U-1)*L < cnt; k += U*L) {
            int v[U], j[U];
            T2 ga[U], gb[U];
            T4 pjv[U], dj[U], gvj[U];
            T rgj[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                v[u] = pj[start+k+u*L];
                j[u] = entryAtom(v[u], bs.ext);
                ga[u] = g12[start+k+u*L];
                if (kStoreG03<T>)
                    gb[u] = g03[start+k+u*L];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                gather(j[u], pjv[u], dj[u]);
                rgj[u] = (kStoreG03<T> && !GVDW) ? T(0) : dj[u].w;
                gvj[u] = GVDW ? gva.gv[j[u]] : dj[u];
            }
            // the U entries of a step summed in T (FP32: one FP64 addition per step instead of per entry), the
            // steps in FP64
            T fu[3] = {-T(0), -T(0), -T(0)}, eu[3] = {-T(0), -T(0), -T(0)}, uu = -T(0);   // -0: folds away
#pragma unroll
            for (int u = 0; u < U; u++)
                finalTerm<T, STRAIN, GVDW>(pi, di, pjv[u], dj[u], ga[u], gb[u], rgi+rgj[u], beta, box, v[u], bs.ext,
                                           acc, fu, eu, uu, gvi, gvj[u], gva, eg);
            acc[0] += (double) uu;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                f[c] += (double) fu[c];
                e[c] += (double) eu[c];
            }
        }
        for (; k < cnt; k += L) {
            const int v = pj[start+k], j = entryAtom(v, bs.ext);
            T2 gb{};
            if (kStoreG03<T>)
                gb = g03[start+k];
            T4 pjr, dj;
            gather(j, pjr, dj);
            const T4 gvj = GVDW ? gva.gv[j] : dj;
            finalTerm<T, STRAIN, GVDW>(pi, di, pjr, dj, g12[start+k], gb, (kStoreG03<T> && !GVDW) ? T(0) : rgi+dj.w,
                                       beta, box, v, bs.ext, acc, f, e, acc[0], gvi, gvj, gva, eg);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            f[c] = groupSum<L>(f[c]);
            e[c] = groupSum<L>(e[c]);
        }
#pragma unroll
        for (int q = 0; q < NACC; q++)
            acc[q] = groupSum<L>(acc[q]);
        if (GVDW)
            eg = groupSum<L>(eg);
        if (lane == 0) {
            force[3*s] = f[0];
            force[3*s+1] = f[1];
            force[3*s+2] = f[2];
            field[3*s] = e[0];
            field[3*s+1] = e[1];
            field[3*s+2] = e[2];
            // each pair is visited from both atoms: half of the sums
#pragma unroll
            for (int q = 0; q < NACC; q++)
                perAtom[NACC*(size_t) s+q] = 0.5*acc[q];
            if (GVDW)
                gva.energy[s] = 0.5*eg;
        }
    }
}

// The GVDW terms of the skipped pairs (exclusions) taken out again after the final pass: per sorted atom, over its
// skipped partners (caller's order CSR), each pair the final pass included (within rc, so in the compacted list,
// and within rvdw) is evaluated by the same code from the same operands and subtracted from the atom's own force,
// GVDW energy and strain rows (no atomics: every visit changes only its own atom's rows).
template<class T, bool STRAIN, bool BATCH>
__global__ void kGvdwSkip(int n, const int* __restrict__ order, const int* __restrict__ rank,
                          const typename Vec4<T>::type* __restrict__ pos, const T* __restrict__ rg,
                          const int* __restrict__ skipStart, const int* __restrict__ skipPartner, BoxSel<T> bs, T rc2,
                          GvdwArgs<T> gva, double* __restrict__ force, double* __restrict__ perAtom) {
    typedef typename Vec4<T>::type T4;
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const int i = order[s];
    const int m0 = skipStart[i], m1 = skipStart[i+1];
    if (m0 == m1)
        return;
    const BoxT<T> box = selectBox<BATCH>(bs, s);
    const T4 pi = pos[s];
    const T4 gvi = gva.gv[s];
    const T rgi = rg[s];
    double eg = 0, f[3] = {0, 0, 0}, acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int m = m0; m < m1; m++) {
        const int sj = rank[skipPartner[m]];
        const T4 pjv = pos[sj];
        T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
        minImage(x, y, z, box);
        const T r2 = x*x+y*y+z*z;
        if (!(r2 < rc2))
            continue;
        gvdwTerm<T, STRAIN>(x, y, z, r2, gvi, gva.gv[sj], rgi+rg[sj], gva.rvdw2, gva.slater != 0, gva.invKe, -1.0, eg,
                            acc, f);
    }
    force[3*s] += f[0];
    force[3*s+1] += f[1];
    force[3*s+2] += f[2];
    gva.energy[s] += 0.5*eg;
    if (STRAIN)
#pragma unroll
        for (int q = 1; q < 10; q++)
            perAtom[10*(size_t) s+q] += 0.5*acc[q];
}

// Parameter derivatives of the pair energy at fixed dipoles d and their tangents along a displacement (x' = g_i - g_j
// for the pair separation x = r_i - r_j) and a dipole change d', per atom i over its compacted pairs (each pair is
// visited from both atoms; each visit takes the part that belongs to its own atom):
//   phi_i = sum_j q_j G0 + (d_j . x) G1                                  (dU/dq_i)
//   E_i   = sum_j q_j x G1 - d_j G1 + (d_j . x) x G2                     (field, -dU/dd_i)
//   R_i   = sum_j dU_ij/da (-2 R_i a^3),  a = (2 (R_i^2 + R_j^2))^-1/2     (dU/dR_i)
// with U_ij = q_i q_j G0 + (q_i d_j.x - q_j d_i.x + d_i.d_j) G1 - (d_i.x)(d_j.x) G2 and
// dG_n/da = (2/sqrt(pi)) (2 a^2)^n exp(-a^2 r^2), the a-derivative of B_n[erf(a r)/r] (the beta part of G_n does
// not depend on a; beyond the damping cut the a part is dropped, and its derivative with it).  The tangents use
// dG_n = -(x . x') G_{n+1}, so G0..G3 stored by the setup pass cover phi' and E'.
template<class T, int L, bool TAN, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kPairTangent(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                           const T* __restrict__ rg, T beta,
                                                           const typename Vec4<T>::type* __restrict__ dip,
                                                           const typename Vec4<T>::type* __restrict__ dipDot,
                                                           const typename Vec4<T>::type* __restrict__ disp,
                                                           const int* __restrict__ candStart,
                                                           const int* __restrict__ nb, const int* __restrict__ pj,
                                                           const typename Vec2<T>::type* __restrict__ g12,
                                                           const typename Vec2<T>::type* __restrict__ g03,
                                                           BoxSel<T> bs, double* __restrict__ out) {
    typedef typename Vec4<T>::type T4;
    typedef typename Vec2<T>::type T2;
    const int gid = xcdBlock()*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const BoxT<T> box = selectBox<BATCH>(bs, s);
    const T4 pi = pos[s];
    const T4 di = dip[s];
    T4 ddi, gi;
    if (TAN) {
        ddi = dipDot[s];
        gi = disp[s];
    }
    const T rgi = rg[s];
    const T Ri = sqrt(T(0.5)*rgi);
    const int start = candStart[s], cnt = nb[s];
    // phi, E[3], R, phi', E'[3], R'
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = lane; k < cnt; k += L) {
        const int ent = pj[start+k], j = entryAtom(ent, bs.ext);
        const T2 ga = g12[start+k];
        const T G1 = ga.x, G2 = ga.y;
        const T4 pjv = pos[j];
        const T4 dj = dip[j];
        T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
        pairImage(x, y, z, box, ent, bs.ext);
        T G0, G3;
        if constexpr (kStoreG03<T>) {
            const T2 gb = g03[start+k];
            G0 = gb.x;
            G3 = gb.y;
        }
        else
            g03FromG12(G1, G2, x*x+y*y+z*z, rgi+rg[j], beta, G0, G3);
        const T qj = pjv.w, qi = pi.w;
        const T djx = dj.x*x+dj.y*y+dj.z*z, dix = di.x*x+di.y*y+di.z*z, didj = di.x*dj.x+di.y*dj.y+di.z*dj.z;
        acc[0] += (double) (qj*G0+djx*G1);
        const T c = qj*G1+djx*G2;
        acc[1] += (double) (c*x-dj.x*G1);
        acc[2] += (double) (c*y-dj.y*G1);
        acc[3] += (double) (c*z-dj.z*G1);
        const T r2 = x*x+y*y+z*z;
        const T a = rsqrtT(rgi+rg[j]);
        const T a2 = a*a;
        const bool damp = a*sqrt(r2) < PairPrec<T>::kDampCut;
        T e0 = 0, e1 = 0, e2 = 0;
        const T w = T(-2)*Ri*a*a2;
        const T mixed = qi*djx-qj*dix+didj;
        if (damp) {
            e0 = T(kTwoOverSqrtPi)*expT(-a2*r2);
            e1 = T(2)*a2*e0;
            e2 = T(2)*a2*e1;
            acc[4] += (double) ((qi*qj*e0+mixed*e1-dix*djx*e2)*w);
        }
        if (TAN) {
            const T4 ddj = dipDot[j];
            const T4 gj = disp[j];
            const T tx = gi.x-gj.x, ty = gi.y-gj.y, tz = gi.z-gj.z;
            const T xx = x*tx+y*ty+z*tz;
            const T djxDot = ddj.x*x+ddj.y*y+ddj.z*z+dj.x*tx+dj.y*ty+dj.z*tz;
            acc[5] += (double) (-qj*xx*G1+djxDot*G1-djx*xx*G2);
            // E' = q_j (x' G1 - x xx G2) - d'_j G1 + d_j xx G2 + djx' x G2 + djx x' G2 - djx x xx G3
            const T cx = qj*G1+djx*G2, cxx = -(qj*G2+djx*G3)*xx+djxDot*G2;
            acc[6] += (double) (cx*tx+cxx*x-ddj.x*G1+dj.x*xx*G2);
            acc[7] += (double) (cx*ty+cxx*y-ddj.y*G1+dj.y*xx*G2);
            acc[8] += (double) (cx*tz+cxx*z-ddj.z*G1+dj.z*xx*G2);
            if (damp) {
                const T dixDot = ddi.x*x+ddi.y*y+ddi.z*z+di.x*tx+di.y*ty+di.z*tz;
                const T didjDot = ddi.x*dj.x+ddi.y*dj.y+ddi.z*dj.z+di.x*ddj.x+di.y*ddj.y+di.z*ddj.z;
                const T sh = T(-2)*a2*xx;   // de_n = sh e_n
                const T mixedDot = qi*djxDot-qj*dixDot+didjDot;
                const T v = (qi*qj*e0+mixed*e1-dix*djx*e2)*sh+mixedDot*e1-(dixDot*djx+dix*djxDot)*e2;
                acc[9] += (double) (v*w);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 10; q++)
        if (TAN || q < 5)
            acc[q] = groupSum<L>(acc[q]);
    if (lane == 0) {
        // out: phi[n], E[3n], R[n], phi'[n], E'[3n], R'[n]
        out[s] = acc[0];
        out[n+3*s] = acc[1];
        out[n+3*s+1] = acc[2];
        out[n+3*s+2] = acc[3];
        out[4*n+s] = acc[4];
        if (TAN) {
            out[5*n+s] = acc[5];
            out[6*n+3*s] = acc[6];
            out[6*n+3*s+1] = acc[7];
            out[6*n+3*s+2] = acc[8];
            out[9*n+s] = acc[9];
        }
    }
}

// The short-range list of the CG preconditioner: one wavefront per atom over its compacted pairs keeps, in
// pair order, those within the preconditioner radius (partner, G1, G2), at most kPreCap; an atom with more sets
// *over, which turns the preconditioner off everywhere (the pairs count whatever the partners' polarizabilities;
// core/cuda/pgm_cuda.cu:1139-1154,1178-1179).
template<class T, bool BATCH>
__global__ void __launch_bounds__(kPairBlock) kShortList(int n, const typename Vec4<T>::type* __restrict__ pos,
                                                         const int* __restrict__ candStart,
                                                         const int* __restrict__ nb, const int* __restrict__ pj,
                                                         const typename Vec2<T>::type* __restrict__ g12,
                                                         BoxSel<T> bs, T rpre2, int* __restrict__ srCount,
                                                         int* __restrict__ srJ,
                                                         typename Vec2<T>::type* __restrict__ srG, int* over) {
    typedef typename Vec4<T>::type T4;
    const int gid = blockIdx.x*blockDim.x+threadIdx.x;
    const int s = gid/64, lane = gid%64;
    if (s >= n)   // uniform per wavefront (64 lanes per atom, blocks of whole wavefronts)
        return;
    const BoxT<T> box = selectBox<BATCH>(bs, s);
    const T4 pi = pos[s];
    const int start = candStart[s], cnt = nb[s];
    const unsigned long long below = (1ull << lane)-1ull;
    int kept = 0;
    for (int k0 = 0; k0 < cnt; k0 += 64) {
        const int k = k0+lane;
        bool in = false;
        int j = 0;
        if (k < cnt) {
            j = pj[start+k];   // the raw entry, kept as is (the preconditioner decodes it)
            const T4 pjv = pos[entryAtom(j, bs.ext)];
            T x = pi.x-pjv.x, y = pi.y-pjv.y, z = pi.z-pjv.z;
            pairImage(x, y, z, box, j, bs.ext);
            in = x*x+y*y+z*z < rpre2;
        }
        const unsigned long long m = __ballot(in);
        if (in) {
            const int q = kept+__popcll(m & below);
            if (q < kPreCap) {
                srJ[(size_t) s*kPreCap+q] = j;
                srG[(size_t) s*kPreCap+q] = g12[start+k];
            }
        }
        kept += __popcll(m);
    }
    if (lane == 0) {
        srCount[s] = min(kept, kPreCap);
        if (kept > kPreCap)
            atomicOr(over, 1);
    }
}

inline int groupsBlocks(int n, int L) { return (int) (((long long) n*L+kPairBlock-1)/kPairBlock); }

// Runs fn with the lane count (16, 32 or 64) and the batch flag as compile-time constants.
template<class Fn>
void dispatchLanes(int lanes, bool batch, Fn fn) {
    auto withBatch = [&](auto L) {
        if (batch)
            fn(L, std::true_type{});
        else
            fn(L, std::false_type{});
    };
    switch (lanes) {
    case 16:
        withBatch(std::integral_constant<int, 16>{});
        break;
    case 32:
        withBatch(std::integral_constant<int, 32>{});
        break;
    case 64:
        withBatch(std::integral_constant<int, 64>{});
        break;
    default:
        throw std::invalid_argument("pgm_hip: pair-pass lanes must be 16, 32 or 64");
    }
}

}   // namespace

template<class T>
void launchPairCompact(hipStream_t st, int n, const typename Vec4<T>::type* pos, const int* candStart,
                       const int* cand, int* nb, int* pj, BoxSel<T> box, T rc2, int lanes) {
    dispatchLanes(lanes, box.table != nullptr, [&](auto L, auto B) {
        kCompact<T, L(), B()><<<groupsBlocks(n, L()), kPairBlock, 0, st>>>(n, pos, candStart, cand, nb, pj, box, rc2);
    });
}

template<class T>
void launchPairSetup(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg,
                     const typename Vec4<T>::type* dip, const typename Vec4<T>::type* dual, const int* candStart,
                     const int* nb, const int* pj, typename Vec2<T>::type* g12, typename Vec2<T>::type* g03,
                     BoxSel<T> box, T beta, double* field, double* fieldDual, int lanes) {
    dispatchLanes(lanes, box.table != nullptr, [&](auto L, auto B) {
        const int blocks = groupsBlocks(n, L());
        if (dual)
            kPairSetup<T, L(), 2, true, B()><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, dip, dual, candStart, nb, pj,
                                                                            g12, g03, box, beta, field, fieldDual);
        else
            kPairSetup<T, L(), 2, false, B()><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, dip, dip, candStart, nb, pj,
                                                                             g12, g03, box, beta, field, nullptr);
    });
}

template<class T>
void launchPairCompactSetup(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg,
                            const typename Vec4<T>::type* dip, const typename Vec4<T>::type* dual,
                            const int* candStart, const int* cand, int* nb, int* pj, typename Vec2<T>::type* g12,
                            typename Vec2<T>::type* g03, BoxSel<T> box, T rc2, T beta, T rgMax, double* field,
                            double* fieldDual, int lanes) {
    dispatchLanes(lanes, box.table != nullptr, [&](auto L, auto B) {
        const int blocks = groupsBlocks(n, L());
        if (dual)
            kCompactSetup<T, L(), true, B()><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, dip, dual, candStart, cand, nb,
                                                                            pj, g12, g03, box, rc2, beta, rgMax,
                                                                            field, fieldDual);
        else
            kCompactSetup<T, L(), false, B()><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, dip, dip, candStart, cand,
                                                                             nb, pj, g12, g03, box, rc2, beta, rgMax,
                                                                             field, nullptr);
    });
}

template<class T>
void launchPairField(hipStream_t st, int n, const typename Vec4<T>::type* pos, const typename Vec4<T>::type* dip,
                     const int* candStart, const int* nb, const int* pj, const typename Vec2<T>::type* g12,
                     BoxSel<T> box, double* field, const int* skip, int lanes, bool otf, T beta, bool reverse) {
    dispatchLanes(lanes, box.table != nullptr, [&](auto L, auto B) {
        if constexpr (sizeof(T) == sizeof(float)) {
            if (otf) {
                kPairFieldOtf<L(), 4, B()><<<groupsBlocks(n, L()), kPairBlock, 0, st>>>(n, pos, dip, candStart, nb,
                                                                                       pj, box, beta, field, skip,
                                                                                       reverse);
                return;
            }
        }
        kPairField<T, L(), 4, B()><<<groupsBlocks(n, L()), kPairBlock, 0, st>>>(n, pos, dip, candStart, nb, pj, g12,
                                                                                box, field, skip, reverse);
    });
}

template<class T>
void launchPairFinal(hipStream_t st, int n, const typename Vec4<T>::type* pos, const typename Vec4<T>::type* dip,
                     const T* rg, T beta, const int* candStart, const int* nb, const int* pj,
                     const typename Vec2<T>::type* g12, const typename Vec2<T>::type* g03, BoxSel<T> box,
                     double* force, double* field, double* perAtom, bool strain, const GvdwArgs<T>& gvdw,
                     int lanes, bool reverse) {
    dispatchLanes(lanes, box.table != nullptr, [&](auto L, auto B) {
        auto go = [&](auto S, auto G) {
            kPairFinal<T, L(), 2, S(), G(), B()><<<groupsBlocks(n, L()), kPairBlock, 0, st>>>(
                n, pos, dip, rg, beta, candStart, nb, pj, g12, g03, box, force, field, perAtom, gvdw, reverse);
        };
        const bool g = gvdw.gv != nullptr;
        if (strain)
            g ? go(std::true_type{}, std::true_type{}) : go(std::true_type{}, std::false_type{});
        else
            g ? go(std::false_type{}, std::true_type{}) : go(std::false_type{}, std::false_type{});
    });
}

template<class T>
void launchGvdwSkip(hipStream_t st, int n, const int* order, const int* rank, const typename Vec4<T>::type* pos,
                    const T* rg, const int* skipStart, const int* skipPartner, BoxSel<T> box, T rc2, bool strain,
                    const GvdwArgs<T>& gvdw, double* force, double* perAtom) {
    const int blocks = (n+255)/256;
    auto go = [&](auto S, auto B) {
        kGvdwSkip<T, S(), B()><<<blocks, 256, 0, st>>>(n, order, rank, pos, rg, skipStart, skipPartner, box, rc2,
                                                        gvdw, force, perAtom);
    };
    const bool b = box.table != nullptr;
    if (strain)
        b ? go(std::true_type{}, std::true_type{}) : go(std::true_type{}, std::false_type{});
    else
        b ? go(std::false_type{}, std::true_type{}) : go(std::false_type{}, std::false_type{});
}

template<class T>
void launchPairTangent(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg, T beta,
                       const typename Vec4<T>::type* dip, const typename Vec4<T>::type* dipDot,
                       const typename Vec4<T>::type* disp, const int* candStart, const int* nb, const int* pj,
                       const typename Vec2<T>::type* g12, const typename Vec2<T>::type* g03, BoxSel<T> box,
                       double* out) {
    const int blocks = groupsBlocks(n, kPairLanes);
    const bool batch = box.table != nullptr;
    if (dipDot && batch)
        kPairTangent<T, kPairLanes, true, true><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, beta, dip, dipDot, disp,
                                                                               candStart, nb, pj, g12, g03, box, out);
    else if (dipDot)
        kPairTangent<T, kPairLanes, true, false><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, beta, dip, dipDot, disp,
                                                                                candStart, nb, pj, g12, g03, box, out);
    else if (batch)
        kPairTangent<T, kPairLanes, false, true><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, beta, dip, dip, pos,
                                                                                candStart, nb, pj, g12, g03, box, out);
    else
        kPairTangent<T, kPairLanes, false, false><<<blocks, kPairBlock, 0, st>>>(n, pos, rg, beta, dip, dip, pos,
                                                                                 candStart, nb, pj, g12, g03, box, out);
}

template<class T>
void launchShortList(hipStream_t st, int n, const typename Vec4<T>::type* pos, const int* candStart, const int* nb,
                     const int* pj, const typename Vec2<T>::type* g12, BoxSel<T> box, T rpre2, int* srCount, int* srJ,
                     typename Vec2<T>::type* srG, int* over) {
    const int blocks = groupsBlocks(n, 64);
    if (box.table != nullptr)
        kShortList<T, true><<<blocks, kPairBlock, 0, st>>>(n, pos, candStart, nb, pj, g12, box, rpre2, srCount, srJ,
                                                          srG, over);
    else
        kShortList<T, false><<<blocks, kPairBlock, 0, st>>>(n, pos, candStart, nb, pj, g12, box, rpre2, srCount, srJ,
                                                           srG, over);
}

#define PGM_INSTANTIATE_PAIRS(T)                                                                                      \
    template void launchShortList<T>(hipStream_t, int, const Vec4<T>::type*, const int*, const int*, const int*,     \
                                     const Vec2<T>::type*, BoxSel<T>, T, int*, int*, Vec2<T>::type*, int*);          \
    template void launchPairCompact<T>(hipStream_t, int, const Vec4<T>::type*, const int*, const int*, int*, int*,   \
                                       BoxSel<T>, T, int);                                                            \
    template void launchPairSetup<T>(hipStream_t, int, const Vec4<T>::type*, const T*, const Vec4<T>::type*,         \
                                     const Vec4<T>::type*, const int*, const int*, const int*, Vec2<T>::type*,        \
                                     Vec2<T>::type*, BoxSel<T>, T, double*, double*, int);                            \
    template void launchPairCompactSetup<T>(hipStream_t, int, const Vec4<T>::type*, const T*, const Vec4<T>::type*,  \
                                            const Vec4<T>::type*, const int*, const int*, int*, int*, Vec2<T>::type*, \
                                            Vec2<T>::type*, BoxSel<T>, T, T, T, double*, double*, int);               \
    template void launchPairField<T>(hipStream_t, int, const Vec4<T>::type*, const Vec4<T>::type*, const int*,       \
                                     const int*, const int*, const Vec2<T>::type*, BoxSel<T>, double*, const int*,    \
                                     int, bool, T, bool);                                                             \
    template void launchPairFinal<T>(hipStream_t, int, const Vec4<T>::type*, const Vec4<T>::type*, const T*, T,      \
                                     const int*, const int*, const int*, const Vec2<T>::type*, const Vec2<T>::type*,  \
                                     BoxSel<T>, double*, double*, double*, bool, const GvdwArgs<T>&, int, bool);      \
    template void launchGvdwSkip<T>(hipStream_t, int, const int*, const int*, const Vec4<T>::type*, const T*,         \
                                    const int*, const int*, BoxSel<T>, T, bool, const GvdwArgs<T>&, double*, double*); \
    template void launchPairTangent<T>(hipStream_t, int, const Vec4<T>::type*, const T*, T, const Vec4<T>::type*,    \
                                       const Vec4<T>::type*, const Vec4<T>::type*, const int*, const int*,            \
                                       const int*, const Vec2<T>::type*, const Vec2<T>::type*, BoxSel<T>, double*);
PGM_INSTANTIATE_PAIRS(float)
PGM_INSTANTIATE_PAIRS(double)

}   // namespace pgmhip
