// Host-side launchers of the pGM HIP kernels (one translation unit per subsystem) and the plain-data
// structures they share with the engine.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace pgmhip {

// -------------------------------------------------------------------------------------------- systems

// The box of an atom.  A batch packs several independent periodic systems into one set of arrays: the atoms
// of a system are contiguous, with the same index range in the caller's and in the sorted order, so one
// system index per atom index serves both orders.  table == nullptr: a single system with box.
// ext != 0: the pair list is the caller's (pgm_set_candidates), whose entries carry their periodic images;
// positions are then used as given, not wrapped into the box.
template<class T>
struct BoxSel {
    BoxT<T> box;
    const BoxT<T>* table;
    const int* sysOf;
    int ext;
};

template<bool BATCH, class T>
__device__ __forceinline__ BoxT<T> selectBox(const BoxSel<T>& b, int idx) {
    if (BATCH)
        return b.table[b.sysOf[idx]];
    return b.box;
}

// An entry of a caller's pair list: partner j in the low kExtIndexBits bits, the image t (|t_k| <= 7 box
// vectors) in the 12 above, code = (t0 + 8) | (t1 + 8) << 4 | (t2 + 8) << 8, for x = r_i - r_j + t.box
// (core/cuda/pgm_cuda.cu:2699-2713 packs the shifts alike).  Entries of the engine's own list are plain j.
constexpr int kExtIndexBits = 20;
constexpr unsigned kExtIndexMask = (1u << kExtIndexBits)-1u;

__host__ __device__ __forceinline__ int extEntry(int j, int t0, int t1, int t2) {
    return (int) ((unsigned) j | ((unsigned) (t0+8) | (unsigned) (t1+8) << 4 | (unsigned) (t2+8) << 8)
                                     << kExtIndexBits);
}

__device__ __forceinline__ int entryAtom(int v, int ext) {
    return ext ? (int) ((unsigned) v & kExtIndexMask) : v;
}

// x = r_i - r_j of a list entry v: the minimum image, or, in a caller's list, plus the entry's image.
template<class T>
__device__ __forceinline__ void pairImage(T& x, T& y, T& z, const BoxT<T>& b, int v, int ext) {
    if (ext) {
        const unsigned c = (unsigned) v >> kExtIndexBits;
        const T t0 = T((int) (c & 15u)-8), t1 = T((int) ((c >> 4) & 15u)-8), t2 = T((int) ((c >> 8) & 15u)-8);
        x += t0*b.ax+t1*b.bx+t2*b.cx;
        y += t1*b.by+t2*b.cy;
        z += t2*b.cz;
    }
    else
        minImage(x, y, z, b);
}

// -------------------------------------------------------------------------------------------- pair list

struct CellGrid {
    int nc[3];     // fractional cells per box dimension
    int ncell;
    int off[3][5];   // distinct neighbour-cell offsets per dimension
    int noff[3];
};

// Cell of every atom (fractional cells of the wrapped position), the per-cell counts and each atom's slot.
void launchCellAssign(hipStream_t st, int n, const double* pos, BoxT<double> box, CellGrid g, int* cellOf, int* slot,
                      int* cellCount, int* subOf);
// Exclusive scan of m ints into out[0..m] (out[m] = total); scratch: >= 1024 ints.
// Exclusive scan of m ints into out[0..m] (out[m] = the total).  Above 4,096 elements it runs over tiles, with
// scratch of scanScratchInts(m) ints (null: one block, slower on long arrays).
int scanScratchInts(int m);
void launchScan(hipStream_t st, const int* in, int* out, int m, int* scratch);
// Sorted order: atoms grouped by cell, inside a cell by (2x2x2 sub-cell, original index).
// base: added to the atom indices written (the system's first atom in a batch).
void launchCellScatter(hipStream_t st, int n, const int* cellOf, const int* slot, const int* cellStart, int* order,
                       int base);
void launchCellSort(hipStream_t st, int ncell, const int* cellStart, const int* subOf, int* order);
void launchRank(hipStream_t st, int n, const int* order, int* rank, int* sortedCell, const int* cellOf, int base);
// Full candidate lists (both directions) of the sorted atoms within rl: counts, then the lists.
void launchCandCount(hipStream_t st, int n, const double4* posS, const int* sortedCell, const int* cellStart,
                     CellGrid g, BoxT<double> box, double rl2, int* count);
void launchCandFill(hipStream_t st, int n, const double4* posS, const int* sortedCell, const int* cellStart,
                    CellGrid g, BoxT<double> box, double rl2, const int* candStart, int* cand, int base);
// Largest squared displacement (minimum image) since the list build, as float bits in *maxBits.
void launchMaxDisplacement(hipStream_t st, int n, const double* pos, const double* posBuild, BoxSel<double> box,
                           unsigned* maxBits);

// -------------------------------------------------------------------------------------------- per atom

// Sorted positions wrapped into the box (xyz, w = charge) in FP64 and, for the mixed path, FP32.
void launchGatherPositions(hipStream_t st, int n, const int* order, const double* pos, const double* qS,
                           BoxSel<double> box, double4* posD, float4* posF);
void launchPermute(hipStream_t st, int n, const int* order, const double* in, double* out);
// The same atoms under new caller indices (newOf: old -> new): order, rank and the build positions renamed.
void launchRename(hipStream_t st, int n, const int* newOf, const int* order, const int* rank, const double* build,
                  int* order2, int* rank2, double* build2);
void launchToFloat(hipStream_t st, int n, const double* in, float* out);
// Caller-order float4 (xyz, w ignored) or float xyz positions to FP64 xyz.
void launchF4ToDouble(hipStream_t st, int n, const float4* posq, double* out);
// *differs = 1 when the count doubles of a and b are not all equal (*differs cleared by the caller).
void launchDiffers(hipStream_t st, long long count, const double* a, const double* b, int* differs);
// Permanent dipoles of the sorted atoms from the covalent CSR (caller's order, positions in FP64).
void launchPermDipoles(hipStream_t st, int n, const int* order, const double* pos, const int* covStart,
                       const int* covPartner, const double* covStrength, BoxSel<double> box, double* perm);

// The per-evaluation preparation of the sorted atoms in one pass: the wrapped positions with charges (FP64
// and, when posF is set, FP32 with the FP32 radius parameters rgF), the permanent dipoles (FP64 and as the
// pair/PME vector type permT, w = the radius parameter), and the clearing of the counters and sums the kernels
// after it accumulate into (zeroI: nZeroI ints, zeroD: nZeroD doubles, maxBits: the next list check's maximum).
struct PrepArgs {
    int n;
    const int* order;
    const double* pos;   // caller's order, FP64 xyz
    const double* qS;    // sorted charges
    const double* rgS;   // sorted radius parameters
    const int* covStart;
    const int* covPartner;
    const double* covStrength;
    BoxSel<double> box;
    double4* posD;
    float4* posF;
    float* rgF;
    double* perm;
    void* permT;
    int* zeroI;
    int nZeroI;
    double* zeroD;
    int nZeroD;
    unsigned* maxBits;
    const double* gvIn;   // GVDW (sqrt A, sqrt C6, b) per atom, caller's order (nullptr: none)
    void* gvOut;          // ... as the pair vector type, sorted order
};
template<class T>
void launchPrepAtoms(hipStream_t st, const PrepArgs& a);

// -------------------------------------------------------------------------------------------- pair passes

// Short-range pairs per atom kept for the CG preconditioner (core/cuda/pgm_cuda.cu:162).
constexpr int kPreCap = 48;

// Lanes per atom of the pair passes: 16, 32 or 64 (picked per engine by the autotuner, 16 by default).
constexpr int kPairLaneChoices[3] = {16, 32, 64};

// Per evaluation: every atom's candidates compacted in place to the pairs within rc (pj: partner; nb: pairs
// per atom), in the atom's candidate segment.
template<class T>
void launchPairCompact(hipStream_t st, int n, const typename Vec4<T>::type* pos, const int* candStart,
                       const int* cand, int* nb, int* pj, BoxSel<T> box, T rc2, int lanes);
// Per evaluation, over the compacted pairs: the kernels (g12: G1, G2; g03: G0, G3, stored by the FP64 path only),
// the direct field of the charges and dipoles dip and, when dual is given, of the dipoles dual alone (the first
// CG product).
template<class T>
void launchPairSetup(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg,
                     const typename Vec4<T>::type* dip, const typename Vec4<T>::type* dual, const int* candStart,
                     const int* nb, const int* pj, typename Vec2<T>::type* g12, typename Vec2<T>::type* g03,
                     BoxSel<T> box, T beta, double* field, double* fieldDual, int lanes);
// launchPairCompact + launchPairSetup in one pass (compacted pairs buffered in LDS per atom); the pairs with
// r^2 < rg_i + rgMax (rgMax: the largest rg, or any value: it orders the segment, nothing else) come last in it.
template<class T>
void launchPairCompactSetup(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg,
                            const typename Vec4<T>::type* dip, const typename Vec4<T>::type* dual,
                            const int* candStart, const int* cand, int* nb, int* pj, typename Vec2<T>::type* g12,
                            typename Vec2<T>::type* g03, BoxSel<T> box, T rc2, T beta, T rgMax, double* field,
                            double* fieldDual, int lanes);
// The dipole part of a CG matrix-vector product from the stored G1, G2 or, otf (FP32 only), with the kernels
// evaluated on the fly from the positions and the radius parameters in the dipoles' w.  reverse (here and in
// launchPairFinal): the atoms in the opposite block order (the same results; the engine alternates the
// direction from one pass over the list to the next, so that a pass starts on what the one before touched last).
template<class T>
void launchPairField(hipStream_t st, int n, const typename Vec4<T>::type* pos, const typename Vec4<T>::type* dip,
                     const int* candStart, const int* nb, const int* pj, const typename Vec2<T>::type* g12,
                     BoxSel<T> box, double* field, const int* skip, int lanes, bool otf = false, T beta = T(0),
                     bool reverse = false);
// The CG preconditioner's short-range list: per atom, its compacted pairs within the radius (srJ partner, srG the
// stored G1, G2; at most kPreCap, counted in srCount); *over set when an atom has more (cleared by the caller).
template<class T>
void launchShortList(hipStream_t st, int n, const typename Vec4<T>::type* pos, const int* candStart, const int* nb,
                     const int* pj, const typename Vec2<T>::type* g12, BoxSel<T> box, T rpre2, int* srCount, int* srJ,
                     typename Vec2<T>::type* srG, int* over);
// The GVDW pair term of the final pass (core/gvdw_kernel.h, combination rule; gv == nullptr: none).
template<class T>
struct GvdwArgs {
    const typename Vec4<T>::type* gv;   // per sorted atom (sqrt A, sqrt C6, b, 0)
    T rvdw2;
    int slater;                         // repulsion: 0 Gaussian, 1 Slater
    T invKe;                            // forces and strain in units of ke, as the electrostatics
    double* energy;                     // per sorted atom: half of its pairs' GVDW energy [kJ/mol]
};

// Energy, forces, field and (strain set) strain of the pairs at the total dipoles from the stored kernels (FP32:
// G0, G3 rebuilt from G1, G2 with the radii rg and beta); per atom (sorted order) perAtom[n][10] = energy, strain
// 0..8 with strain, else perAtom[n] = energy (each pair's half).
template<class T>
void launchPairFinal(hipStream_t st, int n, const typename Vec4<T>::type* pos, const typename Vec4<T>::type* dip,
                     const T* rg, T beta, const int* candStart, const int* nb, const int* pj,
                     const typename Vec2<T>::type* g12, const typename Vec2<T>::type* g03, BoxSel<T> box,
                     double* force, double* field, double* perAtom, bool strain, const GvdwArgs<T>& gvdw,
                     int lanes, bool reverse = false);
// After launchPairFinal with GVDW: the GVDW terms of the skipped pairs (per atom CSR of partners in the caller's
// order, both directions) subtracted from the atoms' forces, GVDW energies and (strain) strain rows.
template<class T>
void launchGvdwSkip(hipStream_t st, int n, const int* order, const int* rank, const typename Vec4<T>::type* pos,
                    const T* rg, const int* skipStart, const int* skipPartner, BoxSel<T> box, T rc2, bool strain,
                    const GvdwArgs<T>& gvdw, double* force, double* perAtom);

// Parameter derivatives of the pair energy at fixed dipoles and, with dipDot (and disp) given, their tangents
// along the displacement disp and the dipole change dipDot, sorted order, each atom from its own visits of its
// pairs.  out (FP64): phi[n], E[3n], dU/dR[n], then with tangents phi'[n], E'[3n], (dU/dR)'[n].
template<class T>
void launchPairTangent(hipStream_t st, int n, const typename Vec4<T>::type* pos, const T* rg, T beta,
                       const typename Vec4<T>::type* dip, const typename Vec4<T>::type* dipDot,
                       const typename Vec4<T>::type* disp, const int* candStart, const int* nb, const int* pj,
                       const typename Vec2<T>::type* g12, const typename Vec2<T>::type* g03, BoxSel<T> box,
                       double* out);

// -------------------------------------------------------------------------------------------- PME

// a / d for 0 <= a < 2^20 and 1 <= d < 2^12 by one multiply: floor(a * ceil(2^32 / d) / 2^32) is exact there
// (the rounding of the reciprocal contributes a * (ceil - 2^32/d) / 2^32 < a * d / 2^32 < 1/d).  Replaces the
// ~20-instruction integer divisions of the FFT index arithmetic.
struct FastDiv {
    unsigned long long magic;
    __host__ __device__ __forceinline__ int div(int a) const {
        return (int) (((unsigned long long) (unsigned) a*magic) >> 32);
    }
};
inline FastDiv fastDiv(int d) { return FastDiv{((1ull << 32)+(unsigned long long) d-1)/(unsigned long long) d}; }

// Stages of a mixed-radix FFT of length K; for each stage the divisions by m = K/R and by Ns = R_0...R_{s-1}.
struct FftPlan {
    int K;
    int nstage;
    int radix[10];
    FastDiv byM[10], byNs[10];
};
FftPlan makePlan(int K);

struct PmeGrid {
    int K[3];
    int K3h;
    int order;
    int brick;     // edge of the spreading bricks in grid points
    int nb[3];     // bricks per dimension
    int nbrick;
    FftPlan plan[3];
};

// One system's PME grid in a batch: its grid, box, the offset of its real grid and its first global brick.
struct PmeSys {
    PmeGrid g;
    BoxT<double> box;
    long long gridOff;
    int brickOff;
    int pad;
};
// The grid of an atom or a brick: one system (g, box), or per system from a table (batch).
struct PmeSel {
    PmeGrid g;
    BoxT<double> box;
    const PmeSys* table;    // batch: per system; nullptr: a single system
    const int* sysOf;       // batch: system of every (sorted) atom index
    const int* brickSys;    // batch: system of every global brick
    int nbrick;             // bricks of all systems
    int order;
    int brick;
    size_t fixedStride;     // FP64 fixed-point grids: the second words' offset (the stacked real grids' size)
};

// Binning of the sorted atoms into spreading bricks and their spline weights, in brick order.
template<class T>
void launchPmeBin(hipStream_t st, int n, const double4* posD, PmeSel ps, int* brickOf, int* slot, int* brickCount);
template<class T>
void launchPmePlace(hipStream_t st, int n, const double4* posD, PmeSel ps, const int* brickOf, const int* slot,
                    const int* brickStart, int* pmeAtom, int* base, T* spl);
// Spread charges (withQ) and dipoles of the sorted atoms onto the fixed-point grids fixed (GridFixed<T>::words
// words per point, the second words of all points after the first words of all points; zeroed by the caller;
// order-independent) or, with dipDot, the change of that density for displacements disp and dipole changes dipDot
// onto the real grids grid (zeroed by the caller).
template<class T>
void launchSpread(hipStream_t st, PmeSel ps, const int* brickStart, const int* pmeAtom, const int* base, const T* spl,
                  const typename Vec4<T>::type* dip, const typename Vec4<T>::type* posq, int withQ, T* grid,
                  long long* fixed, const int* skip, const typename Vec4<T>::type* dipDot = nullptr,
                  const typename Vec4<T>::type* disp = nullptr);
// The PME charge and dipole grids in 64-bit fixed point.  FP32: one word per point, 2^32 units per grid unit, the
// reference's scale (core/cuda/pgm_cuda.cu:52), for |grid values| < 2^31.  FP64: two words per point, the value
// split into hi = rint(v 2^24) and lo = rint((v - hi 2^-24) 2^56): each contribution is rounded at 2^-57 (below
// FP64's own rounding of values of order one), and the sum reads back exactly for |grid values| < 2^29.  The two
// words are accumulated apart (the lo words of up to 2^22 contributions cannot overflow) and combined once.
template<class T>
struct GridFixed;
template<>
struct GridFixed<float> {
    static constexpr int words = 1;
    static constexpr double scale = 4294967296.0, inv = 1.0/4294967296.0;
};
template<>
struct GridFixed<double> {
    static constexpr int words = 2;
    static constexpr double scale = 16777216.0, inv = 1.0/16777216.0;           // hi: 2^24
    static constexpr double loScale = 72057594037927936.0, loInv = 1.0/72057594037927936.0;   // lo: 2^56
};
template<class T>
__device__ __forceinline__ void toFixed(T v, long long& hi, long long& lo) {
    hi = __double2ll_rn((double) v*GridFixed<T>::scale);
    if constexpr (GridFixed<T>::words == 2)
        lo = __double2ll_rn(fma(-(double) hi, GridFixed<T>::inv, (double) v)*GridFixed<T>::loScale);
    else
        lo = 0;
}
// The value of a point whose words are at p[0] and (FP64) p[stride].
template<class T>
__device__ __forceinline__ T fromFixed(const long long* p, size_t stride) {
    if constexpr (GridFixed<T>::words == 2)
        return (T) ((double) p[0]*GridFixed<T>::inv+(double) p[stride]*GridFixed<T>::loInv);
    else
        return (T) ((double) p[0]*GridFixed<T>::inv);
}
// The same, leaving the words zero for the next spread (the forward z transform reads every point once).
template<class T>
__device__ __forceinline__ T takeFixed(long long* p, size_t stride) {
    const T v = fromFixed<T>(p, stride);
    p[0] = 0;
    if constexpr (GridFixed<T>::words == 2)
        p[stride] = 0;
    return v;
}

// One grid shape's share of a merged transform launch: every shape of a batch goes through each stage (z, y,
// x + convolution, y, z) in one launch; a block finds its shape by its block index.
struct FftGroup {
    PmeGrid g;
    int ngrid;                   // stacked grids of this shape
    long long realOff, cplxOff;  // first element of its real grids / complex grids and influence functions
    int twOff;                   // its twiddles
    int partOff;                 // first row of its x-pass partial sums
    int boxOff;                  // its first system's box in the batch box table
    int pz, zy, zx;              // lines per block of the z, y and x stages
    int blk0[4];                 // its first block in the z, y, x and fused z+y launches
};
struct FftLaunch {
    const FftGroup* groups;      // device table
    int ngroups;
    int blocks[4];               // z, y, x and fused z+y launch sizes
    size_t lds[4];               // dynamic LDS of those stages (the largest shape's)
    int zyFused;                 // z and y in one pass per x-plane (many small grids), else apart
};
// The shape of group g's stages (lines per block) and its blocks per stage, as the transforms use them.
template<class T>
void fftGroupShape(FftGroup& G, int blocks[4], size_t lds[4]);
// Forward transforms, influence function (energy and strain partials when asked), inverse transforms of all
// grids of all groups; real: the stacked real grids, cplx/influence/tw: the stacked complex arrays.  boxes: per
// stacked grid (nullptr: box for all).  partials: [.][10] (energy, strain 0..8) per x block, the blocks of a
// group's grid k being rows partOff + [k, k+1) x blocks per grid.  fixed: the forward transforms read the
// fixed-point grids (as the spread left them) instead of real, and leave them zero; the inverse transforms
// write real.
template<class T>
void launchPmeConvolve(hipStream_t st, const FftLaunch& L, T* real, typename Vec2<T>::type* cplx, const T* influence,
                       const typename Vec2<T>::type* tw, BoxT<double> box, const BoxT<double>* boxes, double beta,
                       int energy, int strain, double* partials, const int* skip, long long* fixed,
                       size_t fixedStride);
// Blocks of the x pass per grid (the rows of its energy partials) when ngrid grids of shape g are stacked.
template<class T>
int convolveBlocksPerGrid(PmeGrid g, int ngrid);
// Field -grad phi (sorted order) and, when final, the reciprocal forces.
template<class T>
void launchGather(hipStream_t st, PmeSel ps, int n, const int* pmeAtom, const int* base, const T* spl, const T* phi,
                  const typename Vec4<T>::type* dip, const typename Vec4<T>::type* posq, int final_, double* field,
                  double* force, const int* skip);
// Potential and field of the convolved grid phi at the atoms and, with phiDot, their tangents for atoms displaced
// by disp: out (sorted order) phi[n], E[3n] at offsets 0 and n, phi'[n], E'[3n] at offsets 5n and 6n.
template<class T>
void launchGatherTangent(hipStream_t st, PmeSel ps, int n, const int* pmeAtom, const int* base, const T* spl,
                         const T* phi, const T* phiDot, const typename Vec4<T>::type* disp, double* out);
template<class T>
void launchInfluence(hipStream_t st, PmeGrid g, const double* bmod, BoxT<double> box, double beta, double volume,
                     T* influence);
template<class T>
void launchTwiddles(hipStream_t st, PmeGrid g, typename Vec2<T>::type* tw);
template<class T>
void launchZero(hipStream_t st, T* p, size_t count, const int* skip);

// -------------------------------------------------------------------------------------------- solver

// Scalars of the CG recurrence, device-resident.  done[it] (it = 0..maxIter) is set before iteration it
// runs: every kernel of iteration it returns at once when it is set, so iterations can be enqueued ahead.
struct CgState {
    double norm;      // mean |alpha b| + 1e-300
    double err;       // max |alpha r| / norm
    int iterations;
    int pad;
};

// Partial sums of the CG kernels: regions of kCgMaxBlocks doubles.
constexpr int kCgMaxBlocks = 256;
enum { kPNorm = 0, kPRz0 = 1, kPRz1 = 2, kPpAp = 3, kPBnum = 4, kPMax = 5, kPRegions = 6 };

struct GuessCoef {
    int nc;          // 0: x = alpha b
    int slot[4];
    double c[4];
};

void launchGuess(hipStream_t st, int n, const int* order, const int* gidx, const double* hist, size_t slotStride,
                 GuessCoef gc, const double* alpha, const double* b, double* x);
// b = field_dir + field_rec + cself * perm.
void launchCombineB(hipStream_t st, int n, const double* ed, const double* er, const double* perm, double cself,
                    double* b);
template<class T>
void launchToVec4(hipStream_t st, int n, const double* v, typename Vec4<T>::type* out, const T* rgw = nullptr);
int cgBlocks(int n);
// The CG's preconditioner: Jacobi z = alpha/(1 - alpha cj) r (cj = 0, or the self coefficient), or, with a short
// list (pre), the short-range z = alpha (r + T_s alpha r), computed by launchPrecond after launchCgInit and after
// each launchCgUpdate.
struct Precond {
    bool pre;
    double cj;
};
template<class T>
void launchCgInit(hipStream_t st, int n, const double* alpha, const double* b, const double* x, const double* td,
                  const double* tr, double cself, double tol, int maxIter, double* r, double* z, double* p,
                  typename Vec4<T>::type* pT, double* partials, CgState* cg, int* done, Precond pc,
                  const T* rgw = nullptr);
void launchCgAp(hipStream_t st, int n, int it, const double* alpha, const double* p, const double* td, const double* tr,
                double cself, double* ap, double* partials, const int* done);
void launchCgUpdate(hipStream_t st, int n, int it, const double* alpha, const double* p, const double* ap, double* x,
                    double* r, double* z, double* partials, const int* done, Precond pc);
// z = alpha (r + sum_j T_ij alpha_j r_j) over the short list (partners with alpha_j > 0; M = D when *over), its
// sums into the CG partials: init (after launchCgInit): p = z, r.z; else (after launchCgUpdate of iteration it):
// -step z.Ap and the new r.z (core/cuda/pgm_cuda.cu:1526-1588).
template<class T>
void launchPrecond(hipStream_t st, int n, int it, bool init, const double* alpha, const double* r, const double* ap,
                   const typename Vec4<T>::type* pos, BoxSel<T> box, const int* srCount, const int* srJ,
                   const typename Vec2<T>::type* srG, const int* over, double* z, double* p,
                   typename Vec4<T>::type* pT, double* partials, const int* done, const T* rgw = nullptr);
template<class T>
void launchCgDirection(hipStream_t st, int n, int it, double tol, int maxIter, const double* z, double* p,
                       typename Vec4<T>::type* pT, double* partials, CgState* cg, int* done, const T* rgw = nullptr);
// x += peek alpha r; history push (global order, hist may be NULL); total dipoles d = perm + x, also as the pair
// vector type dT with w = the radius parameter rgS of the (sorted) atom.
template<class T>
void launchPost(hipStream_t st, int n, double peek, const double* alpha, const double* r, double* x, const double* perm,
                const int* order, const int* gidx, double* hist, double* d, const double* rgS,
                typename Vec4<T>::type* dT);
// Total field, and per atom perAtom[n][11]: |d|^2, mu^2/(2 alpha), E_rec (x) d (row-major).
void launchFinalCombine(hipStream_t st, int n, const double* d, const double* mu, const double* alpha,
                        const double* edir, const double* erec, double cself, double* etot, double* perAtom);
// Covalent chain rule per atom (caller's order); per atom strain perAtom[n][9].
void launchCovalentForces(hipStream_t st, int n, const int* rank, const double* pos, const int* covStart,
                          const int* covPartner, const double* covStrength, const int* revStart, const int* revEntry,
                          const int* covAtomOf, const double* etot, BoxSel<double> box, double* fcov,
                          double* perAtom);
// guard (may be null): the forces are added only where *guard is nonzero (a solve's convergence flag).
void launchForcesOut(hipStream_t st, int n, const int* rank, const double* fpair, const double* frec,
                     const double* fcov, double ke, double* fOut, long long* fixedOut, int stride, double fixedScale,
                     const int* guard = nullptr);
// Column sums of partials [nblk][width] into out[0..width).
void launchReducePartials(hipStream_t st, const double* partials, int nblk, int width, double* out);
// Column sums of row ranges: out[outRow(r)*outStride + c] = sum of rows[k][c], k in [starts[r], starts[r+1]),
// for r < nr, c < width; outRow(r) = outIndex[r], or r when outIndex is null.  One block per range.
void launchRangeSums(hipStream_t st, const double* rows, int width, const int* starts, int nr, const int* outIndex,
                     int outStride, double* out);
// In the caller's order: out[i] = v[rank[i]] (3-vectors).
// The operands of a field evaluation of given charges q and dipoles d (caller's order) into the sorted arrays:
// posD.w and, when set, posT.w = q; dS, dT = d (dT.w = rgw).
template<class T>
void launchFieldOperands(hipStream_t st, int n, const int* order, const double* q, const double* d, const T* rgw,
                         double4* posD, typename Vec4<T>::type* posT, double* dS, typename Vec4<T>::type* dT);
void launchUnsort3(hipStream_t st, int n, const int* rank, const double* v, double* out);

// -------------------------------------------------------------------------------------------- parameter gradients

// In the sorted order: out[s] = v[order[s]] (3-vectors).
void launchSort3(hipStream_t st, int n, const int* order, const double* v, double* out);
// Change of the sorted permanent dipoles when the atoms move by disp (caller's order, FP64 xyz).
void launchPermTangent(hipStream_t st, int n, const int* order, const double* pos, const double* disp,
                       const int* covStart, const int* covPartner, const double* covStrength, BoxSel<double> box,
                       double* pdot);
// b = -(E'_pair + E'_rec + cself p') on polarizable rows (pair / PME tangent outputs, E' at offset 6n).
void launchAdjointRhs(hipStream_t st, int n, const double* alpha, const double* pairOut, const double* pmeOut,
                      const double* pdot, double cself, double* b);
// d' = p' - lambda (lambda may be null), FP64 and the pass vector type.
template<class T>
void launchDipoleTangent(hipStream_t st, int n, const double* pdot, const double* lambda, double* ddot,
                         typename Vec4<T>::type* ddotT);

struct ParamAtomArgs {
    int n;
    const int* order;
    const double *q, *alpha, *mu, *lambda, *d, *ddot;   // sorted; lambda, ddot null without a force cotangent
    const double *pairOut, *pmeOut;                     // pair / PME tangent-pass outputs
    double ke, wE, cself;
    double qSelf;                                       // 2 beta / sqrt(pi)
    double qBackground;                                 // pi Q / (V beta^2) (single system)
    const double* qBackSys;                             // batch: pi Q / (V beta^2) per system
    const double* wESys;                                // batch: dL/dE per system
    const int* sysOf;                                   // batch: system of every atom index
    double *gq, *galpha, *gR;                           // caller's order (each may be null)
    double *etot, *etotDot;                             // sorted total field and its tangent
};
// Per-atom charge, polarizability and radius gradients; total field (and tangent) for the covalent terms.
void launchParamAtoms(hipStream_t st, const ParamAtomArgs& a);
// Covalent-strength gradients per entry, written at the entry's input index.
void launchParamCov(hipStream_t st, int ncov, const int* covAtomOf, const int* covPartner, const int* inputOf,
                    const int* rank, const double* pos, const double* disp, const double* etot,
                    const double* etotDot, BoxSel<double> box, double ke, double wE, const double* wESys,
                    double* gc);

}   // namespace pgmhip
