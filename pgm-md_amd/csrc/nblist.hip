// Cell list and full neighbour (candidate) lists of the pGM pair passes, plus the per-atom gathers into the
// spatially sorted order.  Replaces the reference's cellKeys / radix sort / cellRanges / listBuild chain
// (core/cuda/pgm_cuda.cu:445-580, 1930-2002) with a counting sort (atoms inside a cell ordered by 2x2x2
// sub-cell, then atom index: deterministic, and neighbouring list entries spatially close) and sub-wavefront
// candidate scans with ballot compaction into full per-atom lists.
#include "kernels.hpp"

#include <algorithm>

namespace pgmhip {

namespace {

constexpr int kListLanes = 16;   // lanes per atom in the candidate scans

__device__ __forceinline__ void fracCell(double x, double y, double z, const BoxT<double>& b, const CellGrid& g,
                                         int c[3], int& sub) {
    sub = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double t = x*b.recip[0][a]+y*b.recip[1][a]+z*b.recip[2][a];
        t -= floor(t);
        const double w = t*g.nc[a];
        int k = (int) w;
        k = k < 0 ? 0 : (k >= g.nc[a] ? g.nc[a]-1 : k);
        c[a] = k;
        sub |= (w-k >= 0.5 ? 1 : 0) << a;
    }
}

__global__ void kCellAssign(int n, const double* pos, BoxT<double> box, CellGrid g, int* cellOf, int* slot,
                            int* cellCount, int* subOf) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    const bool valid = i < n;   // no early return: the whole wavefront takes part in waveSlot
    int cell = 0;
    if (valid) {
        int c[3], sub;
        fracCell(pos[3*i], pos[3*i+1], pos[3*i+2], box, g, c, sub);
        cell = (c[0]*g.nc[1]+c[1])*g.nc[2]+c[2];
        cellOf[i] = cell;
        subOf[i] = sub;
    }
    const int k = waveSlot(cellCount, cell, valid);
    if (valid)
        slot[i] = k;
}

// Single-block exclusive scan: 1024 threads, each over a contiguous chunk.
__global__ void __launch_bounds__(1024) kScan(const int* in, int* out, int m) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int chunk = (m+1023)/1024;
    const int lo = min(m, t*chunk), hi = min(m, lo+chunk);
    long long s = 0;
    for (int k = lo; k < hi; k++)
        s += in[k];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long v = (t >= o ? part[t-o] : 0);
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t]-s;
    for (int k = lo; k < hi; k++) {
        const int v = in[k];
        out[k] = (int) run;
        run += v;
    }
    if (t == 1023)
        out[m] = (int) part[1023];
}

// Multi-block exclusive scan for long arrays, tiles of kScanTile: (1) every tile's sum, (2) the single-block
// scan above over the tile sums, (3) every tile scanned and offset by its tile's prefix.
constexpr int kScanTile = 1024;   // 256 threads x 4 consecutive elements

__device__ inline int blockExclusiveScan256(int v, int* lds, int& total) {
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int u = t >= o ? lds[t-o] : 0;
        __syncthreads();
        lds[t] += u;
        __syncthreads();
    }
    total = lds[255];
    const int incl = lds[t];
    __syncthreads();
    return incl-v;
}

__global__ void __launch_bounds__(256) kTileSums(const int* in, int m, int* sums) {
    __shared__ int lds[256];
    const int base = blockIdx.x*kScanTile+4*threadIdx.x;
    int s = 0;
    for (int k = 0; k < 4; k++)
        s += base+k < m ? in[base+k] : 0;
    int total;
    (void) blockExclusiveScan256(s, lds, total);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) kTileScan(const int* in, int* out, int m, const int* sumStart) {
    __shared__ int lds[256];
    const int base = blockIdx.x*kScanTile+4*threadIdx.x;
    int v[4], s = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = base+k < m ? in[base+k] : 0;
        s += v[k];
    }
    int total;
    int run = sumStart[blockIdx.x]+blockExclusiveScan256(s, lds, total);
    for (int k = 0; k < 4; k++) {
        if (base+k < m)
            out[base+k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x-1 && threadIdx.x == 0)
        out[m] = sumStart[gridDim.x];
}

__global__ void kCellScatter(int n, const int* cellOf, const int* slot, const int* cellStart, int* order, int base) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        order[cellStart[cellOf[i]]+slot[i]] = i+base;
}

// Inside every cell: ascending (sub-cell, atom index) (the atomic slots above are not deterministic).  One
// wavefront per cell: the cell's keys (unique: the atom index is in them) go to LDS and every key's place is the
// number of keys below it.  A cell of more than kSortCap atoms (far above any physical density at these cell
// sizes) is sorted by insertion in one lane.
constexpr int kSortCap = 512;

__global__ void __launch_bounds__(64) kCellSort(const int* cellStart, const int* subOf, int* order) {
    __shared__ long long key[kSortCap];
    const int c = blockIdx.x, t = threadIdx.x;
    const int lo = cellStart[c], m = cellStart[c+1]-lo;
    if (m <= kSortCap) {
        for (int k = t; k < m; k += 64) {
            const int v = order[lo+k];
            key[k] = ((long long) subOf[v] << 32) | (unsigned) v;
        }
        __syncthreads();
        for (int k = t; k < m; k += 64) {
            const long long kv = key[k];
            int below = 0;
            for (int j = 0; j < m; j++)
                below += key[j] < kv ? 1 : 0;
            order[lo+below] = (int) (unsigned) (kv & 0xffffffffll);
        }
        return;
    }
    if (t != 0)
        return;
    for (int k = lo+1; k < lo+m; k++) {
        const int v = order[k];
        const long long kv = ((long long) subOf[v] << 32) | (unsigned) v;
        int j = k-1;
        while (j >= lo && (((long long) subOf[order[j]] << 32) | (unsigned) order[j]) > kv) {
            order[j+1] = order[j];
            j--;
        }
        order[j+1] = v;
    }
}

__global__ void kRank(int n, const int* order, int* rank, int* sortedCell, const int* cellOf, int base) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const int i = order[s];
    rank[i] = s+base;
    sortedCell[s] = cellOf[i];
}

// One group of L lanes per sorted atom scans the neighbour cells; FILL writes the compacted candidates.
template<bool FILL, int L>
__global__ void __launch_bounds__(256) kCandidates(int n, const double4* posS, const int* sortedCell,
                                                   const int* cellStart, CellGrid g, BoxT<double> box, double rl2,
                                                   int* count, const int* candStart, int* cand, int jOffset) {
    const int gid = xcdBlock()*blockDim.x+threadIdx.x;
    const int s = gid/L, lane = gid%L;
    if (s >= n)
        return;
    const double4 pi = posS[s];
    const int c = sortedCell[s];
    const int c2 = c%g.nc[2], c1 = (c/g.nc[2])%g.nc[1], c0 = c/(g.nc[2]*g.nc[1]);
    const int shift = (threadIdx.x%64)/L*L;
    const unsigned long long below = (lane == 0 ? 0ull : ((1ull << lane)-1ull));
    int found = 0;
    int out = FILL ? candStart[s] : 0;
    const int outEnd = FILL ? candStart[s+1] : 0;
    for (int a = 0; a < g.noff[0]; a++) {
        const int n0 = (c0+g.off[0][a]+g.nc[0])%g.nc[0];
        for (int b = 0; b < g.noff[1]; b++) {
            const int n1 = (c1+g.off[1][b]+g.nc[1])%g.nc[1];
            for (int d = 0; d < g.noff[2]; d++) {
                const int n2 = (c2+g.off[2][d]+g.nc[2])%g.nc[2];
                const int cell = (n0*g.nc[1]+n1)*g.nc[2]+n2;
                const int lo = cellStart[cell], hi = cellStart[cell+1];
                for (int base = lo; base < hi; base += L) {
                    const int j = base+lane;
                    bool in = false;
                    if (j < hi && j != s) {
                        const double4 pj = posS[j];
                        double x = pi.x-pj.x, y = pi.y-pj.y, z = pi.z-pj.z;
                        minImage(x, y, z, box);
                        in = x*x+y*y+z*z < rl2;
                    }
                    if (FILL) {
                        const unsigned long long m = (__ballot(in) >> shift) & (L == 64 ? ~0ull : ((1ull << L)-1ull));
                        const int at = out+__popcll(m & below);
                        if (in && at < outEnd)
                            cand[at] = j+jOffset;
                        out += __popcll(m);
                    }
                    else
                        found += in ? 1 : 0;
                }
            }
        }
    }
    if (!FILL) {
        found = groupSum<L>(found);
        if (lane == 0)
            count[s] = found;
    }
}

template<bool BATCH>
__global__ void kMaxDisplacement(int n, const double* pos, const double* posBuild, BoxSel<double> bs,
                                 unsigned* maxBits) {
    __shared__ double red[16];
    double m = 0;
    for (int i = blockIdx.x*blockDim.x+threadIdx.x; i < n; i += gridDim.x*blockDim.x) {
        const BoxT<double> box = selectBox<BATCH>(bs, i);
        double x = pos[3*i]-posBuild[3*i], y = pos[3*i+1]-posBuild[3*i+1], z = pos[3*i+2]-posBuild[3*i+2];
        minImage(x, y, z, box);
        m = fmax(m, x*x+y*y+z*z);
    }
    m = blockMax(m, red);
    if (threadIdx.x == 0)
        atomicMax(maxBits, __float_as_uint((float) m));
}

__device__ __forceinline__ void wrapPosition(double& x, double& y, double& z, const BoxT<double>& b) {
    const double t0 = floor(x*b.recip[0][0]+y*b.recip[1][0]+z*b.recip[2][0]);
    const double t1 = floor(x*b.recip[0][1]+y*b.recip[1][1]+z*b.recip[2][1]);
    const double t2 = floor(x*b.recip[0][2]+y*b.recip[1][2]+z*b.recip[2][2]);
    x -= t0*b.ax+t1*b.bx+t2*b.cx;
    y -= t1*b.by+t2*b.cy;
    z -= t2*b.cz;
}

template<bool BATCH>
__global__ void kGatherPositions(int n, const int* order, const double* pos, const double* qS, BoxSel<double> bs,
                                 double4* posD, float4* posF) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const BoxT<double> box = selectBox<BATCH>(bs, s);
    const int i = order[s];
    double x = pos[3*i], y = pos[3*i+1], z = pos[3*i+2];
    if (!bs.ext)   // a caller's pair list carries images relative to the positions as given
        wrapPosition(x, y, z, box);
    const double q = qS[s];
    posD[s] = make_double4(x, y, z, q);
    if (posF)
        posF[s] = make_float4((float) x, (float) y, (float) z, (float) q);
}

// The same atoms under new caller indices (newOf: old caller index -> new): the sorted order and the list over
// it stay; order, rank and the positions at the build (caller's order) are renamed into order2, rank2, build2.
__global__ void kRename(int n, const int* newOf, const int* order, const int* rank, const double* build,
                        int* order2, int* rank2, double* build2) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n)
        return;
    const int k = newOf[i];
    order2[i] = newOf[order[i]];   // i read as a sorted index here
    rank2[k] = rank[i];
    build2[3*k] = build[3*i];
    build2[3*k+1] = build[3*i+1];
    build2[3*k+2] = build[3*i+2];
}

__global__ void kPermute(int n, const int* order, const double* in, double* out) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s < n)
        out[s] = in[order[s]];
}

// p_i = sum_m c_m unit(r_partner - r_i), minimum image, FP64 (core/pgm_core.cpp:51-62).
template<bool BATCH>
__global__ void kPermDipoles(int n, const int* order, const double* pos, const int* covStart, const int* covPartner,
                             const double* covStrength, BoxSel<double> bs, double* perm) {
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= n)
        return;
    const BoxT<double> box = selectBox<BATCH>(bs, s);
    const int i = order[s];
    double px = 0, py = 0, pz = 0;
    for (int m = covStart[i]; m < covStart[i+1]; m++) {
        const int j = covPartner[m];
        double x = pos[3*j]-pos[3*i], y = pos[3*j+1]-pos[3*i+1], z = pos[3*j+2]-pos[3*i+2];
        minImage(x, y, z, box);
        const double f = covStrength[m]/sqrt(x*x+y*y+z*z);
        px += f*x;
        py += f*y;
        pz += f*z;
    }
    perm[3*s] = px;
    perm[3*s+1] = py;
    perm[3*s+2] = pz;
}

template<class T, bool BATCH>
__global__ void kPrepAtoms(PrepArgs a) {
    typedef typename Vec4<T>::type T4;
    const int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s < a.nZeroI)
        a.zeroI[s] = 0;
    if (s < a.nZeroD)
        a.zeroD[s] = 0.0;
    if (s < 2 && a.maxBits)   // both words of the next list check: displacement since the build, last move
        a.maxBits[s] = 0u;
    if (s >= a.n)
        return;
    const BoxT<double> box = selectBox<BATCH>(a.box, s);
    const int i = a.order[s];
    const double xi = a.pos[3*i], yi = a.pos[3*i+1], zi = a.pos[3*i+2];
    double x = xi, y = yi, z = zi;
    if (!a.box.ext)
        wrapPosition(x, y, z, box);
    const double q = a.qS[s];
    a.posD[s] = make_double4(x, y, z, q);
    if (a.posF) {
        a.posF[s] = make_float4((float) x, (float) y, (float) z, (float) q);
        a.rgF[s] = (float) a.rgS[s];
    }
    // p_i = sum_m c_m unit(r_partner - r_i), minimum image (core/pgm_core.cpp:51-62)
    double px = 0, py = 0, pz = 0;
    for (int m = a.covStart[i]; m < a.covStart[i+1]; m++) {
        const int j = a.covPartner[m];
        double dx = a.pos[3*j]-xi, dy = a.pos[3*j+1]-yi, dz = a.pos[3*j+2]-zi;
        minImage(dx, dy, dz, box);
        const double f = a.covStrength[m]/sqrt(dx*dx+dy*dy+dz*dz);
        px += f*dx;
        py += f*dy;
        pz += f*dz;
    }
    a.perm[3*s] = px;
    a.perm[3*s+1] = py;
    a.perm[3*s+2] = pz;
    T4 o;
    o.x = (T) px;
    o.y = (T) py;
    o.z = (T) pz;
    o.w = (T) a.rgS[s];   // the radius parameter rides along for the pair passes' partner gathers
    static_cast<T4*>(a.permT)[s] = o;
    if (a.gvIn) {
        T4 g;
        g.x = (T) a.gvIn[3*i];
        g.y = (T) a.gvIn[3*i+1];
        g.z = (T) a.gvIn[3*i+2];
        g.w = T(0);
        static_cast<T4*>(a.gvOut)[s] = g;
    }
}

inline int blocksFor(long long threads, int bs) { return (int) ((threads+bs-1)/bs); }

}   // namespace

template<class T>
void launchPrepAtoms(hipStream_t st, const PrepArgs& a) {
    const int threads = std::max(a.n, std::max(a.nZeroI, a.nZeroD));
    if (a.box.table)
        kPrepAtoms<T, true><<<blocksFor(threads, 256), 256, 0, st>>>(a);
    else
        kPrepAtoms<T, false><<<blocksFor(threads, 256), 256, 0, st>>>(a);
}

template void launchPrepAtoms<float>(hipStream_t, const PrepArgs&);
template void launchPrepAtoms<double>(hipStream_t, const PrepArgs&);

void launchCellAssign(hipStream_t st, int n, const double* pos, BoxT<double> box, CellGrid g, int* cellOf, int* slot,
                      int* cellCount, int* subOf) {
    kCellAssign<<<blocksFor(n, 256), 256, 0, st>>>(n, pos, box, g, cellOf, slot, cellCount, subOf);
}

int scanScratchInts(int m) {
    return m > 4*kScanTile ? 2*((m+kScanTile-1)/kScanTile+1) : 0;
}

void launchScan(hipStream_t st, const int* in, int* out, int m, int* scratch) {
    if (m <= 4*kScanTile || !scratch) {
        kScan<<<1, 1024, 0, st>>>(in, out, m);
        return;
    }
    const int nb = (m+kScanTile-1)/kScanTile;
    int* sums = scratch;
    int* sumStart = scratch+nb+1;
    kTileSums<<<nb, 256, 0, st>>>(in, m, sums);
    kScan<<<1, 1024, 0, st>>>(sums, sumStart, nb);
    kTileScan<<<nb, 256, 0, st>>>(in, out, m, sumStart);
}

void launchCellScatter(hipStream_t st, int n, const int* cellOf, const int* slot, const int* cellStart, int* order,
                       int base) {
    kCellScatter<<<blocksFor(n, 256), 256, 0, st>>>(n, cellOf, slot, cellStart, order, base);
}

void launchCellSort(hipStream_t st, int ncell, const int* cellStart, const int* subOf, int* order) {
    if (ncell > 0)
        kCellSort<<<ncell, 64, 0, st>>>(cellStart, subOf, order);
}

void launchRank(hipStream_t st, int n, const int* order, int* rank, int* sortedCell, const int* cellOf, int base) {
    kRank<<<blocksFor(n, 256), 256, 0, st>>>(n, order, rank, sortedCell, cellOf, base);
}

void launchCandCount(hipStream_t st, int n, const double4* posS, const int* sortedCell, const int* cellStart,
                     CellGrid g, BoxT<double> box, double rl2, int* count) {
    kCandidates<false, kListLanes><<<blocksFor((long long) n*kListLanes, 256), 256, 0, st>>>(
        n, posS, sortedCell, cellStart, g, box, rl2, count, nullptr, nullptr, 0);
}

void launchCandFill(hipStream_t st, int n, const double4* posS, const int* sortedCell, const int* cellStart,
                    CellGrid g, BoxT<double> box, double rl2, const int* candStart, int* cand, int base) {
    kCandidates<true, kListLanes><<<blocksFor((long long) n*kListLanes, 256), 256, 0, st>>>(
        n, posS, sortedCell, cellStart, g, box, rl2, nullptr, candStart, cand, base);
}

void launchMaxDisplacement(hipStream_t st, int n, const double* pos, const double* posBuild, BoxSel<double> box,
                           unsigned* maxBits) {
    const int nb = std::min(1024, blocksFor(n, 256));
    if (box.table)
        kMaxDisplacement<true><<<nb, 256, 0, st>>>(n, pos, posBuild, box, maxBits);
    else
        kMaxDisplacement<false><<<nb, 256, 0, st>>>(n, pos, posBuild, box, maxBits);
}

void launchGatherPositions(hipStream_t st, int n, const int* order, const double* pos, const double* qS,
                           BoxSel<double> box, double4* posD, float4* posF) {
    if (box.table)
        kGatherPositions<true><<<blocksFor(n, 256), 256, 0, st>>>(n, order, pos, qS, box, posD, posF);
    else
        kGatherPositions<false><<<blocksFor(n, 256), 256, 0, st>>>(n, order, pos, qS, box, posD, posF);
}

void launchRename(hipStream_t st, int n, const int* newOf, const int* order, const int* rank, const double* build,
                  int* order2, int* rank2, double* build2) {
    kRename<<<blocksFor(n, 256), 256, 0, st>>>(n, newOf, order, rank, build, order2, rank2, build2);
}

void launchPermute(hipStream_t st, int n, const int* order, const double* in, double* out) {
    kPermute<<<blocksFor(n, 256), 256, 0, st>>>(n, order, in, out);
}

void launchPermDipoles(hipStream_t st, int n, const int* order, const double* pos, const int* covStart,
                       const int* covPartner, const double* covStrength, BoxSel<double> box, double* perm) {
    if (box.table)
        kPermDipoles<true><<<blocksFor(n, 256), 256, 0, st>>>(n, order, pos, covStart, covPartner, covStrength, box,
                                                              perm);
    else
        kPermDipoles<false><<<blocksFor(n, 256), 256, 0, st>>>(n, order, pos, covStart, covPartner, covStrength, box,
                                                               perm);
}

}   // namespace pgmhip

namespace pgmhip {
namespace {
__global__ void kF4ToDouble(int n, const float4* posq, double* out) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n)
        return;
    const float4 p = posq[i];
    out[3*i] = p.x;
    out[3*i+1] = p.y;
    out[3*i+2] = p.z;
}
}   // namespace

void launchF4ToDouble(hipStream_t st, int n, const float4* posq, double* out) {
    kF4ToDouble<<<(n+255)/256, 256, 0, st>>>(n, posq, out);
}

// *differs = 1 when a and b differ anywhere (count doubles; *differs cleared by the caller).  Every writer writes
// the same value, so plain stores suffice.
__global__ void kDiffers(long long count, const double* a, const double* b, int* differs) {
    for (long long i = (long long) blockIdx.x*blockDim.x+threadIdx.x; i < count; i += (long long) gridDim.x*blockDim.x)
        if (a[i] != b[i]) {
            *differs = 1;
            return;
        }
}

void launchDiffers(hipStream_t st, long long count, const double* a, const double* b, int* differs) {
    const int blocks = (int) std::min<long long>(1024, (count+255)/256);
    kDiffers<<<std::max(1, blocks), 256, 0, st>>>(count, a, b, differs);
}
}   // namespace pgmhip

namespace pgmhip {
namespace {
__global__ void kToFloat(int n, const double* in, float* out) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        out[i] = (float) in[i];
}
}   // namespace

void launchToFloat(hipStream_t st, int n, const double* in, float* out) {
    kToFloat<<<(n+255)/256, 256, 0, st>>>(n, in, out);
}
}   // namespace pgmhip
