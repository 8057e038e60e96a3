// pGM engine on one MI355X: host orchestration of the HIP kernels and the C ABI of include/pgm_hip.h.
//
// One evaluation (core/pgm_core.cpp:405-537 on the CPU; core/cuda/pgm_cuda.cu:2172-2334 in the reference's
// GPU engine):
//   positions -> (list check, rebuild if stale) -> sorted positions, permanent dipoles, PME bricks/splines
//   -> pair setup (compacted list + field of q and p) + PME(q, p) -> b
//   -> guess (history) -> CG (matvec = stored-kernel pair pass + PME(dipoles) + self)
//   -> peek, history -> final pair pass + PME with energy/strain + forces gather -> covalent chain rule
//   -> reductions -> terms, forces, strain.
// The CG iterations are enqueued ahead (as many as the last evaluation needed) with per-iteration done
// flags on the device, so a typical evaluation synchronizes with the host only for the list check and the
// result read-back.
//
// A batch (pgm_set_batch) packs several independent periodic systems, each with its own box and PME grid,
// into one engine: every pass is one launch over the atoms (or bricks) of all systems, the transforms run once
// per grid shape over the stacked grids of that shape, and the CG solves the block-diagonal system of all
// of them at once.
#include "kernels.hpp"

#include "../../include/pgm_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace pgmhip {

namespace {

constexpr double kMinPolarizability = 1e-9;   // core/pgm_constants.h:9
constexpr int kNHist = 5;                     // 4 predictor entries + the slot this step writes
// Spreading bricks: the larger edge halves the tiles' halo (17^3 vs 13^3 points flushed per 12^3 vs 8^3 owned,
// 2.8 vs 4.3 atomic additions per grid point at order 6) but makes fewer workgroups; it is used when the engine's
// grids still give kMinBricks of them (4 per CU), the smaller one otherwise.  Measured: the 64-snapshot batch
// (~2,400 bricks of 12) 2,184 -> 2,293 snapshot-gradients/s; the 96k box (216 bricks of 12, 729 of 8) keeps 8,
// its spread being 30% slower with 12.
constexpr int kBrickSmall = 8, kBrickLarge = 12, kMinBricks = 1024;
constexpr int kChunkAtoms = 2048;             // rows per first-stage block of the per-system sums

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("pgm_hip: ")+what+": "+hipGetErrorString(e));
}
#define HIPCHK(x) hipCheck((x), #x)

struct DevBuf;

// A bump allocator over device memory the library does not allocate (the allocation-free entry point's static
// scratch): while one is active on the calling thread, every DevBuf reserve carves from it; the buffers it
// served are given back their own allocations when it ends (release).
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    std::vector<DevBuf*> served;
    void* carve(size_t b);
    void release();
};
thread_local Arena* t_arena = nullptr;

// Device allocation that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    // while an arena serves this buffer: its own allocation, given back by Arena::release
    void* ownP = nullptr;
    size_t ownBytes = 0;
    bool inArena = false;
    ~DevBuf() {
        if (p && !inArena)
            (void) hipFree(p);
        else if (ownP)
            (void) hipFree(ownP);
    }
    void reserve(size_t b) {
        if (t_arena) {
            if (inArena && b <= bytes)
                return;
            if (!inArena) {
                ownP = p;
                ownBytes = bytes;
                inArena = true;
                t_arena->served.push_back(this);
            }
            p = t_arena->carve(std::max<size_t>(b, 256));
            bytes = std::max<size_t>(b, 256);
            return;
        }
        if (b <= bytes)
            return;
        if (p)
            HIPCHK(hipFree(p));
        p = nullptr;
        bytes = 0;
        HIPCHK(hipMalloc(&p, std::max<size_t>(b, 256)));
        bytes = std::max<size_t>(b, 256);
    }
    template<class U>
    U* as() const {
        return static_cast<U*>(p);
    }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
};

void* Arena::carve(size_t b) {
    const size_t at = (used+255) & ~(size_t) 255;
    if (at+b > cap)
        throw std::runtime_error("pgm_hip: the allocation-free evaluation needs more than the " +
                                 std::to_string(cap >> 20)+" MiB of its static scratch (a smaller system, or the "
                                 "allocating entry points)");
    used = at+b;
    return base+at;
}

void Arena::release() {
    for (DevBuf* d : served) {
        d->p = d->ownP;
        d->bytes = d->ownBytes;
        d->ownP = nullptr;
        d->ownBytes = 0;
        d->inArena = false;
    }
    served.clear();
    used = 0;
}

// Pinned host staging that only grows.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~HostBuf() {
        if (p)
            (void) hipHostFree(p);
    }
    void reserve(size_t b) {
        if (b <= bytes)
            return;
        if (t_arena)
            throw std::logic_error("pgm_hip: the allocation-free evaluation would allocate pinned host memory");
        if (p)
            HIPCHK(hipHostFree(p));
        p = nullptr;
        HIPCHK(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
    }
    template<class U>
    U* as() const {
        return static_cast<U*>(p);
    }
};


// The scratch of the allocation-free entry point (pgm_compute_device_f64_noalloc): a static device array,
// one per GPU the library runs on, held until the process ends.  Host code uses its address from
// hipGetSymbolAddress (scratchAddress), never the array's name, which is only a host-side stand-in.
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kNoAllocMaxAtoms = 4096;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratchMutex;
std::vector<void*> g_scratchPtr;   // per device ordinal

// The device address of the scratch on the current device (looked up once per device, at engine creation).
void* scratchAddress() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if ((int) g_scratchPtr.size() <= dev)
        g_scratchPtr.resize(dev+1, nullptr);
    if (!g_scratchPtr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratchPtr[dev] = ptr;
    }
    return g_scratchPtr[dev];
}

// The PME stream of the engines on the current device, shared by them and destroyed with the last one.  HIP
// deals ordinary streams out over a small pool of hardware queues (GPU_MAX_HW_QUEUES, 4 by default); two
// streams on one queue run one after the other, and an engine's two streams did land on one queue once a
// process held more than four streams.  A stream created with a CU mask gets a hardware queue of its own, so
// this one is, with every CU in the mask (bits past the device's CUs are ignored).  It synchronizes with the
// null stream, as a stream created without flags does.
std::mutex g_pmeStreamMutex;
std::vector<hipStream_t> g_pmeStream;   // per device ordinal
std::vector<int> g_pmeStreamUsers;

hipStream_t acquirePmeStream() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_pmeStreamMutex);
    if ((int) g_pmeStream.size() <= dev) {
        g_pmeStream.resize(dev+1, nullptr);
        g_pmeStreamUsers.resize(dev+1, 0);
    }
    if (!g_pmeStream[dev]) {
        std::vector<uint32_t> mask(32, 0xffffffffu);   // 1,024 CUs
        if (hipExtStreamCreateWithCUMask(&g_pmeStream[dev], (uint32_t) mask.size(), mask.data()) != hipSuccess) {
            (void) hipGetLastError();   // an ordinary stream then (it may share a queue with the engine's)
            g_pmeStream[dev] = nullptr;
            HIPCHK(hipStreamCreateWithFlags(&g_pmeStream[dev], hipStreamNonBlocking));
        }
    }
    g_pmeStreamUsers[dev]++;
    return g_pmeStream[dev];
}

// with the device current, after the caller's work on the stream has finished
void releasePmeStream() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return;
    std::lock_guard<std::mutex> lock(g_pmeStreamMutex);
    if (dev >= (int) g_pmeStream.size() || !g_pmeStream[dev] || --g_pmeStreamUsers[dev] > 0)
        return;
    (void) hipStreamDestroy(g_pmeStream[dev]);
    g_pmeStream[dev] = nullptr;
}

template<class T>
BoxT<T> makeBox(const double H[3][3], const double recip[3][3]) {
    BoxT<T> b;
    b.ax = (T) H[0][0];
    b.bx = (T) H[1][0];
    b.by = (T) H[1][1];
    b.cx = (T) H[2][0];
    b.cy = (T) H[2][1];
    b.cz = (T) H[2][2];
    b.iax = (T) (1.0/H[0][0]);
    b.iby = (T) (1.0/H[1][1]);
    b.icz = (T) (1.0/H[2][2]);
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++)
            b.recip[k][a] = (T) recip[k][a];
    return b;
}

// |sum_k M_p(k+1) e^{2 pi i m k/K}|^2 with zeros replaced by the mean of their neighbours
// (core/pgm_math.h:98-126).
std::vector<double> bsplineModuli(int K, int p) {
    std::vector<double> prev(p+1, 0.0), cur(p+1, 0.0);
    prev[0] = 1.0;
    for (int k = 2; k <= p; k++) {
        for (int j = 0; j < k; j++) {
            const double hi = (j < k-1 ? prev[j] : 0.0), lo = (j > 0 ? prev[j-1] : 0.0);
            cur[j] = (j*hi+(k-j)*lo)/(k-1);
        }
        for (int j = 0; j < k; j++)
            prev[j] = cur[j];
    }
    std::vector<double> mod(K);
    for (int m = 0; m < K; m++) {
        double re = 0, im = 0;
        for (int k = 0; k+1 < p && k < K; k++) {
            const double ang = 2*kPi*m*k/K;
            re += prev[k+1]*std::cos(ang);
            im += prev[k+1]*std::sin(ang);
        }
        mod[m] = re*re+im*im;
    }
    for (int m = 0; m < K; m++)
        if (mod[m] < 1e-7)
            mod[m] = 0.5*(mod[(m+K-1)%K]+mod[(m+1)%K]);
    return mod;
}


// The result slots of one system's evaluation (device, read back once).
enum {
    kResPair = 0,      // energy, strain[9] (pair, already halved)
    kResRec = 10,      // energy, strain[9] (k-space)
    kResComb = 20,     // |d|^2, polarization, E_rec (x) d [9]
    kResCov = 31,      // covalent strain [9]
    kResGvdw = 40,     // GVDW energy
    kResCount = 41,
};

}   // namespace

namespace {

enum class PosSource { HostF64, HostF32, DeviceF64, DeviceF4 };

// Kernel groups timed with HIP events when timing is on (pgm_set_timing).
enum Phase {
    kPhList = 0,      // list check and rebuild
    kPhPrep,          // sorted positions, permanent dipoles, PME bricks and splines
    kPhPairCompact,   // candidates within the cutoff, compacted in place
    kPhPairSetup,     // stored kernels + field of q and p (and the first CG product)
    kPhPairField,     // stored-kernel dipole field (one per matrix-vector product)
    kPhSpread,
    kPhFft,           // z, y, x+convolution, y, z transforms
    kPhGather,
    kPhCg,            // guess, CG vector kernels, post
    kPhPairFinal,
    kPhFinal,         // total field, covalent chain rule, reductions, forces out
    kNumPhases
};

class DeviceEngine {
public:
    explicit DeviceEngine(const pgm_settings& s) : s_(s) {
        validate();
        withDevice([&] {
            HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&evIn_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evOut_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evSync_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evList_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evFork_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evJoin_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evDone_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evCov_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evTab_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evParam_, hipEventDisableTiming));
            pmeSt_ = st_;
            // pinned words read back by the evaluations (allocated here: a launch makes no allocation that
            // could wait for the GPU)
            listH_.reserve(64);
            cgH_.reserve(64);
            scalH_.reserve(64);
            resH_.reserve(kResCount*sizeof(double));   // one system's sums
            scratchPtr_ = scratchAddress();
            st2_ = acquirePmeStream();   // last: released by the destructor, which runs only for a built engine
        });
    }
    ~DeviceEngine() {
        int prev = -1;
        (void) hipGetDevice(&prev);
        if (s_.device >= 0)
            (void) hipSetDevice(s_.device);
        if (st_)
            (void) hipStreamSynchronize(st_);
        for (hipEvent_t ev : evPool_)
            (void) hipEventDestroy(ev);
        if (evIn_)
            (void) hipEventDestroy(evIn_);
        if (evOut_)
            (void) hipEventDestroy(evOut_);
        if (st2_) {
            (void) hipStreamSynchronize(st2_);
            releasePmeStream();
        }
        for (hipEvent_t ev : {evSync_, evList_, evFork_, evJoin_, evDone_, evCov_, evTab_, evParam_})
            if (ev)
                (void) hipEventDestroy(ev);
        if (st_)
            (void) hipStreamDestroy(st_);
        if (prev >= 0)
            (void) hipSetDevice(prev);
    }

    std::string lastError;

    template<class Fn>
    void withDevice(Fn fn) {
        int prev = -1;
        if (s_.device >= 0) {
            HIPCHK(hipGetDevice(&prev));
            if (prev != s_.device)
                HIPCHK(hipSetDevice(s_.device));
        }
        try {
            fn();
        }
        catch (...) {
            if (prev >= 0 && prev != s_.device)
                (void) hipSetDevice(prev);
            throw;
        }
        if (prev >= 0 && prev != s_.device)
            HIPCHK(hipSetDevice(prev));
    }

    void setAtoms(int n, const double* q, const double* radius, const double* alpha, const int* gidx, int nGlobal) {
        if (n <= 0 || !q || !radius || !alpha)
            throw std::invalid_argument("pgm_hip: setAtoms needs n > 0 and q, radius, alpha");
        if (nGlobal < n && gidx)
            throw std::invalid_argument("pgm_hip: setAtoms: n_global < n");
        const int ng = gidx ? nGlobal : n;
        std::vector<int>& gi = giTmp_;   // members, so that a fitting step's call allocates nothing on the host
        gi.resize(n);
        bool identity = true;
        for (int i = 0; i < n; i++) {
            gi[i] = gidx ? gidx[i] : i;
            identity = identity && gi[i] == i;
            if (gi[i] < 0 || gi[i] >= ng)
                throw std::invalid_argument("pgm_hip: setAtoms: global index out of range");
            if (!(radius[i] > 0))
                throw std::invalid_argument("pgm_hip: setAtoms: radii must be positive");
        }
        withDevice([&] {
            // q, the radius parameter 2R^2 and alpha into pinned staging before the stream is drained: a fitting
            // step sets them at every step, and this loop then runs while the GPU finishes the step before.  The
            // staging is free once its last copies have run.
            HIPCHK(hipEventSynchronize(evParam_));
            paramH_.reserve(3*sizeof(double)*n);
            double* stage = paramH_.as<double>();
            std::memcpy(stage, q, n*sizeof(double));
            bool polarizable = false;
            double rgMax = 0.0;
            for (int i = 0; i < n; i++) {
                stage[n+i] = 2.0*radius[i]*radius[i];
                rgMax = std::max(rgMax, stage[n+i]);
                const double a = alpha[i] > kMinPolarizability ? alpha[i] : 0.0;
                stage[2*n+i] = a;
                polarizable = polarizable || a > 0;
            }
            HIPCHK(hipStreamSynchronize(st_));
            const bool resized = n != n_ || ng != nGlobal_;
            if (resized) {
                histSlots_.clear();
                haveList_ = false;
                lastIterations_ = 0;
            }
            // The same atoms again (new parameters, or another order named by the global indices): the list over
            // the spatially sorted atoms is kept and renamed (core/cuda/pgm_cuda.cu:2519-2533,2586-2632).
            // newOf[old caller index] = new caller index of the same global atom; the same identity order as
            // before (new parameters only) needs no map.
            std::vector<int>& newOf = newOfTmp_;
            newOf.clear();
            const bool keep = !resized && haveList_ && !extList_ && (int) gidxH_.size() == n;
            const bool sameOrder = keep && identity && gidxIdentity_;
            if (keep && !sameOrder) {
                std::vector<int> localOf(ng, -1);
                for (int i = 0; i < n; i++)
                    localOf[gi[i]] = i;
                newOf.resize(n);
                std::vector<char> used(n, 0);
                std::vector<int> sysOfAtom;
                if (batch_) {   // a permutation must keep every atom in its system of a batch
                    sysOfAtom.resize(n);
                    for (size_t y = 0; y < sys_.size(); y++)
                        for (int i = sys_[y].a0; i < sys_[y].a1; i++)
                            sysOfAtom[i] = (int) y;
                }
                for (int i = 0; i < n && !newOf.empty(); i++) {
                    const int k = localOf[gidxH_[i]];
                    if (k < 0 || used[k] || (batch_ && k != i && sysOfAtom[k] != sysOfAtom[i]))
                        newOf.clear();
                    else {
                        used[k] = 1;
                        newOf[i] = k;
                    }
                }
            }
            if (n != n_ || sys_.empty()) {
                // a new atom count: one system (a batch partition is set again with pgm_set_batch)
                sys_.assign(1, Sys{});
                sys_[0].a0 = 0;
                sys_[0].a1 = n;
                for (int a = 0; a < 3; a++)
                    sys_[0].grid[a] = s_.grid[a];
                batch_ = false;
                layoutDirty_ = true;
            }
            n_ = n;
            nGlobal_ = ng;
            qH_.assign(q, q+n);
            refreshCharges();
            polarizable_ = polarizable;
            rgMax_ = rgMax;
            // the staged parameters copied without a wait: every kernel that reads them is enqueued on the same
            // stream after the copies
            DevBuf* dst[3] = {&qD_, &rgD_, &alphaD_};
            for (int k = 0; k < 3; k++) {
                dst[k]->reserve(n*sizeof(double));
                HIPCHK(hipMemcpyAsync(dst[k]->p, stage+k*n, n*sizeof(double), hipMemcpyHostToDevice, st_));
            }
            HIPCHK(hipEventRecord(evParam_, st_));
            if (resized || !(identity && gidxIdentity_) || gidxH_ != gi)
                upload(gidxD_, gi.data(), n*sizeof(int));
            gidxH_.swap(gi);
            gidxIdentity_ = identity;
            if (sameOrder)
                permuteParameters();
            else if (newOf.empty())
                haveList_ = false;
            else
                renameList(newOf);
            gvdwOn_ = false;     // GVDW is set again after every setAtoms
            extList_ = false;    // and so is a caller's pair list
            if (covN_ < 0)
                setCovalentImpl(0, nullptr, nullptr, nullptr);
            haveResults_ = false;
        });
    }

    // GVDW by the combination rule over the pairs within rvdw except the skipped ones (core/cuda/pgm_cuda.h:57-65);
    // rvdw 0: off.  The pair list's radius is the electrostatic cutoff, so rvdw may not exceed it.
    void setGvdw(int repulsion, double rvdw, const double* sqrtA, const double* sqrtC6, const double* b, int nskip,
                 const int* skipI, const int* skipJ) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: setGvdw before setAtoms");
        if (extList_ && rvdw > 0)
            throw std::invalid_argument("pgm_hip: GVDW with a caller's pair list (setCandidates) is not supported");
        if (!(rvdw > 0)) {
            gvdwOn_ = false;
            return;
        }
        if (repulsion != 0 && repulsion != 1)
            throw std::invalid_argument("pgm_hip: setGvdw: repulsion must be 0 (Gaussian) or 1 (Slater)");
        if (rvdw > s_.cutoff)
            throw std::invalid_argument("pgm_hip: setGvdw: rvdw larger than the cutoff is not supported (the pair "
                                        "list radius is the electrostatic cutoff)");
        if (!sqrtA || !sqrtC6 || !b || nskip < 0 || (nskip > 0 && (!skipI || !skipJ)))
            throw std::invalid_argument("pgm_hip: setGvdw: bad arrays");
        std::vector<int> start(n_+1, 0), part(2*(size_t) nskip);
        for (int k = 0; k < nskip; k++) {
            const int i = skipI[k], j = skipJ[k];
            if (i < 0 || i >= n_ || j < 0 || j >= n_ || i == j)
                throw std::invalid_argument("pgm_hip: setGvdw: skipped pair index out of range");
            if (systemOf(i) != systemOf(j))
                throw std::invalid_argument("pgm_hip: setGvdw: a skipped pair joins two systems of the batch");
            start[i+1]++;
            start[j+1]++;
        }
        for (int i = 0; i < n_; i++)
            start[i+1] += start[i];
        std::vector<int> fill(start.begin(), start.end()-1);
        for (int k = 0; k < nskip; k++) {
            part[fill[skipI[k]]++] = skipJ[k];
            part[fill[skipJ[k]]++] = skipI[k];
        }
        std::vector<double> gv(3*(size_t) n_);
        for (int i = 0; i < n_; i++) {
            gv[3*i] = sqrtA[i];
            gv[3*i+1] = sqrtC6[i];
            gv[3*i+2] = b[i];
        }
        withDevice([&] {
            HIPCHK(hipStreamSynchronize(st_));
            upload(gvInD_, gv.data(), gv.size()*sizeof(double));
            upload(skipStartD_, start.data(), start.size()*sizeof(int));
            if (part.empty())
                skipPartD_.reserve(sizeof(int));
            else
                upload(skipPartD_, part.data(), part.size()*sizeof(int));
        });
        nSkip_ = nskip;
        rvdw_ = rvdw;
        gvdwSlater_ = repulsion == 1;
        gvdwOn_ = true;
        gvdwReady_ = false;
        haveResults_ = false;
    }

    // The caller's pair candidates instead of the engine's own list, until the next setAtoms
    // (Engine::setCandidates, core/cuda/pgm_cuda.h:70-73, core/cuda/pgm_cuda.cu:2694-2737): segment s holds the
    // pairs (segI[s], candJ[k]), k in [segBegin[s], segBegin[s+1]), with x = r_i + t.box - r_j,
    // t = segShift[3s..3s+2] integer box vectors (|t_k| <= 7); every pair at most once; those beyond the cutoff
    // are skipped at every evaluation.  Both directions go into per-atom segments in the caller's order (no
    // spatial sort), each entry packing its image (extEntry).  Positions are then used as given, not wrapped.
    void setCandidates(int nseg, const int* segI, const int* segShift, const int* segBegin, const int* candJ) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: setCandidates before setAtoms");
        if (batch_)
            throw std::invalid_argument("pgm_hip: setCandidates: one system only (not with set_batch)");
        if (gvdwOn_)
            throw std::invalid_argument("pgm_hip: GVDW with a caller's pair list (setCandidates) is not supported");
        if (n_ > (int) kExtIndexMask+1)
            throw std::invalid_argument("pgm_hip: setCandidates: at most 2^20 atoms");
        if (nseg < 0 || (nseg > 0 && (!segI || !segShift || !segBegin)))
            throw std::invalid_argument("pgm_hip: setCandidates: missing arrays");
        const int n = n_;
        std::vector<int> count(n+1, 0);
        for (int sg = 0; sg < nseg; sg++) {
            const int i = segI[sg];
            if (i < 0 || i >= n)
                throw std::invalid_argument("pgm_hip: setCandidates: atom index out of range");
            if (segBegin[sg+1] < segBegin[sg] || segBegin[sg] < 0)
                throw std::invalid_argument("pgm_hip: setCandidates: segment bounds not increasing");
            for (int a = 0; a < 3; a++)
                if (std::abs(segShift[3*sg+a]) > 7)
                    throw std::invalid_argument("pgm_hip: setCandidates: periodic shift beyond 7 box vectors");
            if (segBegin[sg+1] > segBegin[sg] && !candJ)
                throw std::invalid_argument("pgm_hip: setCandidates: missing arrays");
            for (int k = segBegin[sg]; k < segBegin[sg+1]; k++) {
                const int j = candJ[k];
                if (j < 0 || j >= n)
                    throw std::invalid_argument("pgm_hip: setCandidates: atom index out of range");
                if (j == i && segShift[3*sg] == 0 && segShift[3*sg+1] == 0 && segShift[3*sg+2] == 0)
                    throw std::invalid_argument("pgm_hip: setCandidates: an atom paired with itself");
                count[i+1]++;
                count[j+1]++;
            }
        }
        for (int i = 0; i < n; i++)
            count[i+1] += count[i];
        std::vector<int> fill(count.begin(), count.end()-1), cand(std::max(count[n], 1));
        for (int sg = 0; sg < nseg; sg++) {
            const int i = segI[sg], *t = segShift+3*sg;
            for (int k = segBegin[sg]; k < segBegin[sg+1]; k++) {
                const int j = candJ[k];
                cand[fill[i]++] = extEntry(j, t[0], t[1], t[2]);
                cand[fill[j]++] = extEntry(i, -t[0], -t[1], -t[2]);
            }
        }
        extStartH_.swap(count);
        extCandH_.swap(cand);
        extList_ = true;
        haveList_ = false;
        haveResults_ = false;
    }

    // CG preconditioner (options of the reference's CUDA engine, core/cuda/pgm_cuda.cu:1753-1760): the short-range
    // M = D + D T_s D over the pairs within radius (0: Jacobi), at most kPreCap per atom or Jacobi everywhere;
    // Jacobi with the self term of the diagonal, alpha/(1 - alpha c_self).  Kept across setAtoms.
    void setPreconditioner(double radius, bool jacobiSelf) {
        if (!(radius >= 0) || radius > s_.cutoff)
            throw std::invalid_argument("pgm_hip: the preconditioner radius must be in [0, cutoff]");
        preRadius_ = radius;
        jacobiSelf_ = jacobiSelf;
        preReady_ = false;
        haveResults_ = false;
    }

    void setCovalent(int ncov, const int* atom, const int* partner, const double* strength) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: setCovalent before setAtoms");
        if (ncov < 0 || (ncov > 0 && (!atom || !partner || !strength)))
            throw std::invalid_argument("pgm_hip: setCovalent: bad arrays");
        // the same entries with new strengths (a fitting step): only the strengths change on the device
        if (ncov == covN_ && covBuiltN_ == n_ && (int) covEntryOfH_.size() == ncov &&
            std::equal(atom, atom+ncov, covAtomH_.begin()) && std::equal(partner, partner+ncov, covPartnerH_.begin())) {
            // through pinned staging, copied without a wait: the copy is ordered on the stream before every
            // kernel that reads the strengths, and the staging is written again only after its last copy
            withDevice([&] {
                HIPCHK(hipEventSynchronize(evCov_));
                covH_.reserve(std::max(ncov, 1)*sizeof(double));
                double* str = covH_.as<double>();
                for (int m = 0; m < ncov; m++)
                    str[covEntryOfH_[m]] = strength[m];
                if (ncov)
                    HIPCHK(hipMemcpyAsync(covStrengthD_.p, str, ncov*sizeof(double), hipMemcpyHostToDevice, st_));
                HIPCHK(hipEventRecord(evCov_, st_));
            });
            haveResults_ = false;
            return;
        }
        std::vector<int> sysOfAtom;
        if (sys_.size() > 1) {
            sysOfAtom.resize(n_);
            for (size_t y = 0; y < sys_.size(); y++)
                for (int i = sys_[y].a0; i < sys_[y].a1; i++)
                    sysOfAtom[i] = (int) y;
        }
        for (int m = 0; m < ncov; m++) {
            if (atom[m] < 0 || atom[m] >= n_ || partner[m] < 0 || partner[m] >= n_ || atom[m] == partner[m])
                throw std::invalid_argument("pgm_hip: setCovalent: atom index out of range");
            if (!sysOfAtom.empty() && sysOfAtom[atom[m]] != sysOfAtom[partner[m]])
                throw std::invalid_argument("pgm_hip: setCovalent: a covalent dipole joins two systems of the batch");
        }
        withDevice([&] {
            HIPCHK(hipStreamSynchronize(st_));
            setCovalentImpl(ncov, atom, partner, strength);
        });
    }

    // Partition the atoms into k independent periodic systems: atoms [offsets[s], offsets[s+1]) form system s
    // with PME grid grids[3s..3s+2] (nullptr: the settings' grid for every system).
    void setSystems(int k, const int* offsets, const int* grids) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: set_batch before setAtoms");
        if (k < 1 || !offsets)
            throw std::invalid_argument("pgm_hip: set_batch needs k >= 1 and k+1 offsets");
        if (offsets[0] != 0 || offsets[k] != n_)
            throw std::invalid_argument("pgm_hip: set_batch: offsets must run from 0 to the atom count");
        for (int s = 0; s < k; s++) {
            if (offsets[s+1] <= offsets[s])
                throw std::invalid_argument("pgm_hip: set_batch: every system needs at least one atom");
            for (int a = 0; a < 3; a++) {
                const int K = grids ? grids[3*s+a] : s_.grid[a];
                if (K < s_.order)
                    throw std::invalid_argument("pgm_hip: set_batch: a PME grid is smaller than the order");
                (void) makePlan(K);
            }
        }
        for (size_t m = 0; m < covAtomH_.size(); m++) {
            const int sa = (int) (std::upper_bound(offsets, offsets+k+1, covAtomH_[m])-offsets)-1;
            const int sp = (int) (std::upper_bound(offsets, offsets+k+1, covPartnerH_[m])-offsets)-1;
            if (sa != sp)
                throw std::invalid_argument("pgm_hip: set_batch: a covalent dipole joins two systems");
        }
        withDevice([&] {
            HIPCHK(hipStreamSynchronize(st_));
            sys_.assign(k, Sys{});
            for (int s = 0; s < k; s++) {
                sys_[s].a0 = offsets[s];
                sys_[s].a1 = offsets[s+1];
                for (int a = 0; a < 3; a++)
                    sys_[s].grid[a] = grids ? grids[3*s+a] : s_.grid[a];
            }
            batch_ = k > 1;
            gvdwOn_ = false;   // GVDW (its skipped pairs are checked per system) is set again after set_batch
            extList_ = false;
            layoutDirty_ = true;
            haveList_ = false;
            haveResults_ = false;
            refreshCharges();
        });
    }
    int numSystems() const { return (int) sys_.size(); }

    // One evaluation of every system.  boxes: 9 doubles per system.  fHost / fDev / fixed: force outputs
    // (added); strainOut (single system only) added to; terms: one per system.
    void evaluate(PosSource src, const void* pos, const double* boxes, bool newStep, double* fHost, double* fDev,
                  long long* fixed, int stride, double fixedScale, double* strainOut, pgm_terms* terms,
                  hipStream_t caller, bool keepForces) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: compute before setAtoms");
        if (!pos || !boxes)
            throw std::invalid_argument("pgm_hip: compute needs positions and a box");
        if (strainOut && sys_.size() > 1)
            throw std::invalid_argument("pgm_hip: the strain derivative is per system: not available for a batch");
        adjointReady_ = false;   // set again only when this evaluation completes
        prevValid_ = false;      // (a next launch's move prediction starts over)
        withDevice([&] {
            HIPCHK(hipEventRecord(evIn_, caller));
            HIPCHK(hipStreamWaitEvent(st_, evIn_, 0));
            setBoxes(boxes);
            stagePositions(src, pos);
            // the list check's round trip overlaps the preparation kernels; a stale list (rare) is rebuilt
            // and the atoms prepared again in the new order
            const bool checking = haveList_ && !extList_;   // a caller's list is the caller's to keep
            tb(kPhList);
            if (checking)
                startListCheck();
            else if (!haveList_)
                buildList();
            te();
            const bool mixed = s_.precision == PGM_MIXED;
            mixed ? prep<float>() : prep<double>();
            if (checking && listStale()) {
                tb(kPhList);
                buildList();
                te();
                mixed ? prep<float>() : prep<double>();
            }
            // Once the engine has solved before, the CG is enqueued ahead as a launch's is (the most iterations
            // the last 16 solves needed, one more if they varied, each skipped on the device once converged): no
            // host wait between the solve and the final stage.  The forces are added only if those iterations
            // sufficed (the solve's device flag); a solve that needed more is evaluated again, from the same guess
            // (the history is advanced only by a complete solve), with its convergence checked on the host.
            bool ahead = polarizable_ && s_.polarizable && numRecent_ > 0;
            for (;;) {
                asyncEval_ = asyncCg_ = ahead;
                asyncPredict_ = predictedIterations();
                asyncStop_ = -1;
                pendingPush_ = -1;
                try {
                    if (mixed)
                        run<float>(newStep, strainOut != nullptr);
                    else
                        run<double>(newStep, strainOut != nullptr);
                }
                catch (...) {
                    asyncEval_ = asyncCg_ = false;
                    throw;
                }
                asyncEval_ = asyncCg_ = false;
                const int* guard = (ahead && asyncStop_ >= 0 && asyncStop_ < s_.max_iter)
                                       ? doneD_.as<int>()+asyncStop_ : nullptr;
                // forces out, caller's order
                tb(kPhFinal);
                if (fDev || fixed)
                    launchForcesOut(st_, n_, rankD_.as<int>(), fpairD_.as<double>(), frecD_.as<double>(),
                                    fcovD_.as<double>(), s_.ke, fDev, fixed, stride, fixedScale, guard);
                if (fHost || keepForces) {
                    forcesOutD_.reserve(3*sizeof(double)*n_);
                    launchZero<double>(st_, forcesOutD_.as<double>(), 3*(size_t) n_, nullptr);
                    launchForcesOut(st_, n_, rankD_.as<int>(), fpairD_.as<double>(), frecD_.as<double>(),
                                    fcovD_.as<double>(), s_.ke, forcesOutD_.as<double>(), nullptr, 0, 0.0);
                    stageH_.reserve(3*sizeof(double)*n_);
                    HIPCHK(hipMemcpyAsync(stageH_.p, forcesOutD_.p, 3*sizeof(double)*n_, hipMemcpyDeviceToHost,
                                          st_));
                }
                te();
                // the evaluation's positions, for a backward at the same positions (which then skips evaluating)
                lastPosD_.reserve(3*sizeof(double)*n_);
                HIPCHK(hipMemcpyAsync(lastPosD_.p, posDev_, 3*sizeof(double)*n_, hipMemcpyDeviceToDevice, st_));
                const size_t resBytes = sys_.size()*kResCount*sizeof(double);
                resH_.reserve(resBytes);
                HIPCHK(hipMemcpyAsync(resH_.p, resD_.p, resBytes, hipMemcpyDeviceToHost, st_));
                waitStream();
                if (!ahead || asyncStop_ < 0)
                    break;
                CgState cs;
                int done;
                std::memcpy(&cs, cgH_.p, sizeof(CgState));
                std::memcpy(&done, cgH_.as<char>()+sizeof(CgState), sizeof(int));
                const bool shortSolve = !done && asyncStop_ < s_.max_iter;
                noteIterations(shortSolve ? asyncStop_+1 : cs.iterations);
                if (!shortSolve) {
                    commitPush(pendingPush_);
                    iterations_ = lastIterations_ = cs.iterations;
                    residual_ = cs.err;
                    break;
                }
                ahead = false;
                shortSolves_++;
                timed_.clear();   // (the brackets of the pass given up have all run: the wait above)
                evUsed_ = 0;
            }
            pendingPush_ = -1;
            asyncStop_ = -1;
            lastBoxes_.assign(boxes, boxes+9*sys_.size());
            adjointReady_ = true;
            collectTiming();
            assemble(strainOut != nullptr);
            if (fHost) {
                const double* f = stageH_.as<double>();
                for (int k = 0; k < 3*n_; k++)
                    fHost[k] += f[k];
            }
            if (keepForces)
                forcesH_.assign(stageH_.as<double>(), stageH_.as<double>()+3*n_);
            if (strainOut)
                for (int k = 0; k < 9; k++)
                    strainOut[k] += sys_[0].strain[k];
            if (terms)
                for (size_t s = 0; s < sys_.size(); s++)
                    terms[s] = sys_[s].terms;
            HIPCHK(hipEventRecord(evOut_, st_));
            HIPCHK(hipStreamWaitEvent(caller, evOut_, 0));
        });
    }

    // launchDevice (core/cuda/pgm_cuda.cu:2754-2774): the whole evaluation enqueued behind the caller's stream, no
    // host wait.  The pair list is checked on the device and read by the completion; before the launch the host
    // predicts from the last step's displacement and largest move whether the list will hold and rebuilds it if
    // not (core/cuda/pgm_cuda.cu:1912-1928).  The CG iterations enqueued are the most that the last 16 solves
    // needed (one more if their counts varied), each behind its device flag; the first launch of an engine (no
    // such count yet), and a repeat after a solve that needed more, wait on the host between chunks of iterations
    // instead.  The forces are added to
    // the fixed-point buffer in any case; the completion says whether they are valid.
    void launchDevice(const void* posq4, const double* boxes, bool newStep, long long* fixed, int stride,
                      double fixedScale, hipStream_t caller) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: launch before setAtoms");
        if (!posq4 || !boxes || !fixed)
            throw std::invalid_argument("pgm_hip: launch needs positions, a box and a force buffer");
        settle();
        adjointReady_ = false;
        withDevice([&] {
            HIPCHK(hipEventRecord(evIn_, caller));
            HIPCHK(hipStreamWaitEvent(st_, evIn_, 0));
            setBoxes(boxes);
            stagePositions(PosSource::DeviceF4, posq4);
            // (a repeat is at the positions of the launch it repeats: no move ahead to predict, and none to
            // measure, so the last step's move is kept for the next launch)
            const bool repeat = repeatNext_;
            const double ahead = repeat ? 0.0 : 2.0*std::sqrt(lastS2_);
            repeatNext_ = false;
            bool built = false;
            if (!haveList_ || (!extList_ && staleFor(lastD2_, ahead))) {
                tb(kPhList);
                buildList();
                te();
                built = true;
            }
            pendingChecking_ = !extList_ && !built;
            pendingPrev_ = pendingChecking_ && prevValid_ && !repeat;
            if (pendingChecking_)
                startListCheck(pendingPrev_);
            posPrevD_.reserve(3*sizeof(double)*n_);
            HIPCHK(hipMemcpyAsync(posPrevD_.p, posDev_, 3*sizeof(double)*n_, hipMemcpyDeviceToDevice, st_));
            prevValid_ = !extList_;
            const bool mixed = s_.precision == PGM_MIXED;
            mixed ? prep<float>() : prep<double>();
            asyncEval_ = true;
            asyncCg_ = !syncCgNext_ && numRecent_ > 0;
            asyncPredict_ = predictedIterations();
            syncCgNext_ = false;
            pendingPush_ = -1;
            try {
                mixed ? run<float>(newStep, false) : run<double>(newStep, false);
            }
            catch (...) {
                asyncEval_ = asyncCg_ = false;
                throw;
            }
            const bool cgAhead = asyncCg_;
            asyncEval_ = asyncCg_ = false;
            if (!cgAhead)
                asyncStop_ = -1;   // the solve completed on the host: iterations_, residual_ are final
            tb(kPhFinal);
            launchForcesOut(st_, n_, rankD_.as<int>(), fpairD_.as<double>(), frecD_.as<double>(), fcovD_.as<double>(),
                            s_.ke, nullptr, fixed, stride, fixedScale);
            te();
            const size_t resBytes = sys_.size()*kResCount*sizeof(double);
            resH_.reserve(resBytes);
            HIPCHK(hipMemcpyAsync(resH_.p, resD_.p, resBytes, hipMemcpyDeviceToHost, st_));
            HIPCHK(hipEventRecord(evDone_, st_));
            inFlight_ = true;
        });
    }

    // finishDevice (core/cuda/pgm_cuda.cu:2782-2788,2932-2959): waits for the launch; false when its pair list
    // was stale for those positions (the list is rebuilt now) or its solve needed more iterations than were
    // enqueued: the forces added are not valid, and the caller repeats the launch (same positions, same guess
    // and history entry, which the completion has not advanced).
    bool finishDevice() {
        if (!inFlight_)
            return true;
        bool valid = true;
        withDevice([&] { valid = complete(true); });
        return valid;
    }

    // waitDevice (core/cuda/pgm_cuda.cu:2777-2780): the caller's stream waits on the GPU for the launch's forces.
    void waitDevice(hipStream_t stream) {
        withDevice([&] {
            if (inFlight_)
                HIPCHK(hipStreamWaitEvent(stream, evDone_, 0));
            else
                waitOn(stream);
        });
    }

    bool launchPending() const {
        return inFlight_ && hipEventQuery(evDone_) == hipErrorNotReady;
    }

    // The completion of a launch that nobody finished (waitDevice): at the next call of any kind.  Its forces
    // were used, so a stale list or a short solve found now only counts as late (lateListChecks).
    void settle() {
        if (!inFlight_)
            return;
        withDevice([&] { (void) complete(false); });
    }

    bool complete(bool canRepeat) {
        inFlight_ = false;
        waitEvent(evDone_);
        bool stale = false;
        if (pendingChecking_) {
            float d2, s2;
            std::memcpy(&d2, listH_.p, sizeof(float));
            std::memcpy(&s2, listH_.as<char>()+sizeof(float), sizeof(float));
            if (pendingPrev_)
                lastS2_ = s2;
            stale = staleFor(d2, 0.0);
            lastD2_ = stale ? 0.0 : d2;
        }
        else
            lastD2_ = 0.0;   // a list built for this launch
        int its = iterations_;
        double err = residual_;
        bool shortSolve = false;
        if (asyncStop_ >= 0 && polarizable_ && s_.polarizable) {
            CgState cs;
            int done;
            std::memcpy(&cs, cgH_.p, sizeof(CgState));
            std::memcpy(&done, cgH_.as<char>()+sizeof(CgState), sizeof(int));
            its = cs.iterations;
            err = cs.err;
            shortSolve = !done && asyncStop_ < s_.max_iter;
            noteIterations(shortSolve ? asyncStop_+1 : its);
        }
        if (stale) {
            tb(kPhList);
            buildList();
            te();
            lastD2_ = 0.0;
        }
        if (stale || shortSolve) {
            if (canRepeat) {
                repeats_++;
                repeatNext_ = true;
                syncCgNext_ = shortSolve;
                collectTiming();
                return false;
            }
            lateChecks_++;
        }
        commitPush(pendingPush_);
        pendingPush_ = -1;
        iterations_ = its;
        lastIterations_ = its;
        residual_ = err;
        collectTiming();
        assemble(false);
        return true;
    }

    // pgm_compute_device_f64 without any allocation: every device buffer the evaluation reserves is carved from
    // the library's static scratch of this device, and nothing of it outlives the call: the pair list is built,
    // the solve starts from alpha E (the history is neither read nor written), and the engine's next call of any
    // kind sets up its own buffers again.  Calls on one GPU share its scratch: calls that overlap race.
    void evaluateNoAlloc(const double* pos, const double* box, double* fDev, double* strainOut, pgm_terms* terms,
                         hipStream_t caller) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: compute before setAtoms");
        if (n_ > kNoAllocMaxAtoms)
            throw std::invalid_argument("pgm_hip: the allocation-free evaluation takes at most " +
                                        std::to_string(kNoAllocMaxAtoms)+" atoms (this engine has "+
                                        std::to_string(n_)+")");
        if (extList_)
            throw std::invalid_argument("pgm_hip: the allocation-free evaluation builds its own pair list");
        Arena arena;
        arena.base = static_cast<char*>(scratchPtr_);
        arena.cap = kScratchBytes;
        const bool wasTiming = timing_;
        auto restore = [&] {
            t_arena = nullptr;
            arena.release();
            timing_ = wasTiming;
            noHistory_ = false;
            forgetDeviceState();
        };
        forgetDeviceState();
        timing_ = false;
        noHistory_ = true;
        t_arena = &arena;
        try {
            evaluate(PosSource::DeviceF64, pos, box, true, nullptr, fDev, nullptr, 0, 0.0, strainOut, terms, caller,
                     false);
        }
        catch (...) {
            restore();
            throw;
        }
        restore();
    }
    // Everything derived on the device from positions and parameters is to be set up again (the buffers that
    // held it were the scratch's).
    void forgetDeviceState() {
        haveList_ = false;
        layoutDirty_ = true;
        tablesDirty_ = true;
        for (Sys& y : sys_)
            y.haveInfluence = false;
        haveResults_ = false;
        adjointReady_ = false;
        gvdwReady_ = false;
        preReady_ = false;
        prevValid_ = false;
        fixedClean_ = nullptr;
        fixedCleanWords_ = 0;
    }

    // pgmcore::Engine::setGeometry + field (core/pgm_core.h:76-79, core/pgm_core.cpp:539-560): the field [e/nm^2]
    // at every atom of the charges q and dipoles d given here (caller's order; the engine's own charges, covalent
    // and induced dipoles play no part), direct + reciprocal + self, at positions pos and box: the pair list checked
    // (or built), the atoms prepared, then the setup pair pass and a PME solve of (q, d) as an evaluation's field of
    // the permanent multipoles.  out: 3n, caller's order, assigned.
    void field(const double* pos, const double* boxes, const double* q, const double* d, double* out) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: field before setAtoms");
        if (!pos || !boxes || !q || !d || !out)
            throw std::invalid_argument("pgm_hip: field needs positions, a box, charges, dipoles and an output");
        adjointReady_ = false;
        haveResults_ = false;
        prevValid_ = false;
        withDevice([&] {
            HIPCHK(hipStreamSynchronize(st_));
            setBoxes(boxes);
            stagePositions(PosSource::HostF64, pos);
            const bool mixed = s_.precision == PGM_MIXED;
            if (haveList_ && !extList_) {
                startListCheck();
                if (listStale())
                    buildList();
            }
            else if (!haveList_)
                buildList();
            mixed ? fieldOf<float>(q, d, out) : fieldOf<double>(q, d, out);
        });
    }

    template<class T>
    void fieldOf(const double* q, const double* d, double* out) {
        typedef typename Vec4<T>::type T4;
        const int n = n_;
        prep<T>();
        fqD_.reserve(n*sizeof(double));
        fdD_.reserve(3*n*sizeof(double));
        HIPCHK(hipMemcpyAsync(fqD_.p, q, n*sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipMemcpyAsync(fdD_.p, d, 3*n*sizeof(double), hipMemcpyHostToDevice, st_));
        const bool f32 = sizeof(T) == sizeof(float);
        launchFieldOperands<T>(st_, n, orderD_.as<int>(), fqD_.as<double>(), fdD_.as<double>(), rgT<T>(),
                               posDD_.as<double4>(), f32 ? posT<T>() : nullptr, permD_.as<double>(), permTD_.as<T4>());
        for (DevBuf* b : {&edD_, &erD_, &bD_, &tdD_})
            b->reserve(3*sizeof(double)*n);
        nbD_.reserve(n*sizeof(int));
        pairSetup<T>(setupLanes(), fused_, nullptr);
        pmeSolve<T>(permTD_.as<T4>(), true, posT<T>(), 0, false, erD_.as<double>(), nullptr, nullptr, nullptr);
        const double cself = 4.0*s_.beta*s_.beta*s_.beta/(3.0*kSqrtPi);
        launchCombineB(st_, n, edD_.as<double>(), erD_.as<double>(), permD_.as<double>(), cself, bD_.as<double>());
        tmpD_.reserve(3*sizeof(double)*n);
        launchUnsort3(st_, n, rankD_.as<int>(), bD_.as<double>(), tmpD_.as<double>());
        HIPCHK(hipMemcpyAsync(out, tmpD_.p, 3*sizeof(double)*n, hipMemcpyDeviceToHost, st_));
        HIPCHK(hipStreamSynchronize(st_));
    }

    // True when an evaluation finished at bitwise these positions and boxes and nothing has changed the engine's
    // state since (parameters, settings, the pair passes' timing).  Compares on the device, ordered after the
    // caller's stream; one small read back.
    bool sameAsLastEvaluation(PosSource src, const void* pos, const double* boxes, hipStream_t caller) {
        if (!haveResults_ || !adjointReady_ || lastBoxes_.size() != 9*sys_.size() ||
            !std::equal(lastBoxes_.begin(), lastBoxes_.end(), boxes))
            return false;
        int differs = 1;
        withDevice([&] {
            HIPCHK(hipEventRecord(evIn_, caller));
            HIPCHK(hipStreamWaitEvent(st_, evIn_, 0));
            const size_t bytes = 3*sizeof(double)*n_;
            const double* cand = nullptr;
            if (src == PosSource::DeviceF64)
                cand = static_cast<const double*>(pos);
            else {
                cmpPosD_.reserve(bytes);
                if (src == PosSource::DeviceF4)
                    launchF4ToDouble(st_, n_, static_cast<const float4*>(pos), cmpPosD_.as<double>());
                else {
                    stageH_.reserve(bytes);
                    double* d = stageH_.as<double>();
                    if (src == PosSource::HostF64)
                        std::memcpy(d, pos, bytes);
                    else
                        for (int k = 0; k < 3*n_; k++)
                            d[k] = static_cast<const float*>(pos)[k];
                    HIPCHK(hipMemcpyAsync(cmpPosD_.p, stageH_.p, bytes, hipMemcpyHostToDevice, st_));
                }
                cand = cmpPosD_.as<double>();
            }
            cmpFlagD_.reserve(sizeof(int));
            launchZero<int>(st_, cmpFlagD_.as<int>(), 1, nullptr);
            launchDiffers(st_, 3LL*n_, cand, lastPosD_.as<double>(), cmpFlagD_.as<int>());
            cmpH_.reserve(64);
            HIPCHK(hipMemcpyAsync(cmpH_.p, cmpFlagD_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
            waitStream();
            std::memcpy(&differs, cmpH_.p, sizeof(int));
        });
        return differs == 0;
    }

    // Parameter gradients of a loss L(E, F) at the given positions: evaluates there (same step: the last solution
    // as the guess) unless the last evaluation was at exactly these positions and boxes with nothing changed
    // since, then the adjoint.  g: dL/dF (caller's order, host or device, may be null), wE: dL/dE per system.
    void backward(PosSource src, const void* pos, const double* boxes, const double* g, bool gDevice,
                  const double* wE, double* gq, double* gR, double* galpha, double* gc, bool outDevice,
                  hipStream_t caller) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: backward before setAtoms");
        if (gvdwOn_)
            throw std::invalid_argument("pgm_hip: backward covers the electrostatics only: set GVDW off (rvdw 0)");
        if (sameAsLastEvaluation(src, pos, boxes, caller))
            posDev_ = lastPosD_.as<double>();   // the adjoint reads the evaluation's positions from the engine's copy
        else
            evaluate(src, pos, boxes, false, nullptr, nullptr, nullptr, 0, 0.0, nullptr, nullptr, caller, false);
        adjointReady_ = false;   // the adjoint's solve reuses the CG buffers: a next backward evaluates again
        withDevice([&] {
            HIPCHK(hipEventRecord(evIn_, caller));
            HIPCHK(hipStreamWaitEvent(st_, evIn_, 0));
            const size_t v3 = 3*sizeof(double)*n_;
            const int ncov = std::max(covN_, 0);
            const double* gDev = nullptr;
            if (g && gDevice)
                gDev = g;
            else if (g) {
                gInD_.reserve(v3);
                stageH_.reserve(v3);
                std::memcpy(stageH_.p, g, v3);
                HIPCHK(hipMemcpyAsync(gInD_.p, stageH_.p, v3, hipMemcpyHostToDevice, st_));
                gDev = gInD_.as<double>();
            }
            double *oq = gq, *oR = gR, *oa = galpha, *oc = gc;
            if (!outDevice) {
                gradD_.reserve((3*(size_t) n_+ncov+1)*sizeof(double));
                double* base = gradD_.as<double>();
                oq = gq ? base : nullptr;
                oR = gR ? base+n_ : nullptr;
                oa = galpha ? base+2*n_ : nullptr;
                oc = gc ? base+3*n_ : nullptr;
            }
            // the adjoint's kernels stay out of the phase timing of the evaluations
            const bool wasTiming = timing_;
            timing_ = false;
            try {
                if (s_.precision == PGM_MIXED)
                    adjoint<float>(gDev, wE, oq, oR, oa, oc);
                else
                    adjoint<double>(gDev, wE, oq, oR, oa, oc);
            }
            catch (...) {
                timing_ = wasTiming;
                throw;
            }
            timing_ = wasTiming;
            HIPCHK(hipStreamSynchronize(st_));
            if (!outDevice) {
                const size_t bytes = (3*(size_t) n_+ncov)*sizeof(double);
                stageH_.reserve(bytes);
                HIPCHK(hipMemcpyAsync(stageH_.p, gradD_.p, bytes, hipMemcpyDeviceToHost, st_));
                HIPCHK(hipStreamSynchronize(st_));
                const double* h = stageH_.as<double>();
                if (gq)
                    std::copy(h, h+n_, gq);
                if (gR)
                    std::copy(h+n_, h+2*n_, gR);
                if (galpha)
                    std::copy(h+2*n_, h+3*n_, galpha);
                if (gc)
                    std::copy(h+3*n_, h+3*n_+ncov, gc);
            }
            HIPCHK(hipEventRecord(evOut_, st_));
            HIPCHK(hipStreamWaitEvent(caller, evOut_, 0));
        });
    }
    int adjointIterations() const { return adjIterations_; }
    double adjointResidual() const { return adjResidual_; }

    // Times the pair passes at every lane count on the current lists and keeps the fastest (per-evaluation cost:
    // compaction + setup + (iterations + 1) field passes + final pass).  Returns the lanes chosen.
    int autotune(int reps) {
        if (!haveResults_)
            throw std::logic_error("pgm_hip: autotune needs an evaluation first (the lists of the workload)");
        if (reps < 1)
            throw std::invalid_argument("pgm_hip: autotune needs reps >= 1");
        int best = lanes_;
        bool bestFused = fused_, bestOtf = fieldOtf_;
        withDevice([&] {
            double bestMs = 1e300;
            const bool mixed = s_.precision == PGM_MIXED;
            for (int v = 0; v < 2; v++) {
                const bool fused = v == 0;
                for (int c = 0; c < 3; c++) {
                    const int lanes = kPairLaneChoices[c];
                    tuneMs_[3*v+c] = 1e300;
                    // FP32 also tries the field pass with its kernels evaluated on the fly
                    for (int o = 0; o < (mixed ? 2 : 1); o++) {
                        const bool otf = o == 1;
                        const double ms = mixed ? timePairPasses<float>(lanes, fused, reps, otf)
                                                : timePairPasses<double>(lanes, fused, reps, false);
                        tuneMs_[3*v+c] = std::min(tuneMs_[3*v+c], ms);
                        if (ms < bestMs) {
                            bestMs = ms;
                            best = lanes;
                            bestFused = fused;
                            bestOtf = otf;
                        }
                    }
                }
            }
            lanes_ = best;
            fused_ = bestFused;
            fieldOtf_ = bestOtf;
            // the setup pass on its own: its best lane count need not be the other passes' (it tests every
            // candidate, they stream the compacted pairs)
            double bestSetup = 1e300;
            setupLanes_ = 0;
            for (int c = 0; c < 3; c++) {
                const int lanes = kPairLaneChoices[c];
                const double ms = mixed ? timeSetup<float>(lanes, fused_, reps)
                                        : timeSetup<double>(lanes, fused_, reps);
                if (ms < bestSetup) {
                    bestSetup = ms;
                    setupLanes_ = lanes;
                }
            }
            // then the PME solves beside the pair passes or after them on the engine stream, timed over a CG
            // product (field pass + PME solve) with the chosen pair passes.  Beside, the two share the CUs and
            // the L2; at 999k atoms in FP64, where each fills the chip, that was the slower way.
            for (int o = 0; o < 2; o++)
                overlapMs_[o] = mixed ? timeProduct<float>(o == 1, reps) : timeProduct<double>(o == 1, reps);
            pmeOverlap_ = overlapMs_[1] <= overlapMs_[0];
        });
        return best;
    }
    bool pmeOverlap() const { return pmeOverlap_; }
    void setPmeOverlap(bool on) { pmeOverlap_ = on; }
    const double* overlapTimes() const { return overlapMs_; }
    bool fieldOtf() const { return fieldOtf_; }
    void setFieldOtf(bool on) {
        if (on && s_.precision != PGM_MIXED)
            throw std::invalid_argument("pgm_hip: the on-the-fly field pass is an FP32 (PGM_MIXED) variant");
        fieldOtf_ = on;
    }
    int lanes() const { return lanes_; }
    // the setup pass's lanes: the autotuner's own choice for it, else the other passes'
    int setupLanes() const { return setupLanes_ ? setupLanes_ : lanes_; }
    void setSetupLanes(int lanes) {
        if (lanes != 16 && lanes != 32 && lanes != 64)
            throw std::invalid_argument("pgm_hip: pair-pass lanes must be 16, 32 or 64");
        setupLanes_ = lanes;
    }
    bool fused() const { return fused_; }
    void setLanes(int lanes, bool fused) {
        if (lanes != 16 && lanes != 32 && lanes != 64)
            throw std::invalid_argument("pgm_hip: pair-pass lanes must be 16, 32 or 64");
        lanes_ = lanes;
        setupLanes_ = 0;
        fused_ = fused;
    }
    const double* tuneTimes() const { return tuneMs_; }

    int iterations() const { return iterations_; }
    double residual() const { return residual_; }
    long listBuilds() const { return listBuilds_; }
    long listRenames() const { return listRenames_; }
    long long numCandidates() const { return nCand_; }
    // Pairs within the cutoff at the last evaluation (each counted once).
    long long numPairs() {
        if (!haveResults_)
            return 0;
        long long total = 0;
        withDevice([&] {
            std::vector<int> nb(n_);
            HIPCHK(hipMemcpyAsync(nb.data(), nbD_.p, n_*sizeof(int), hipMemcpyDeviceToHost, st_));
            HIPCHK(hipStreamSynchronize(st_));
            for (int v : nb)
                total += v;
        });
        return total/2;
    }
    long lateListChecks() const { return lateChecks_; }
    long shortSolves() const { return shortSolves_; }
    // the launch-time list prediction's inputs: displacement since the build at the last completed check, the
    // largest move of the last step, the smallest skin; the launches repeated, the synchronous evaluations redone
    void listCheckState(double* out) const {
        double skin = 1e300;
        for (const Sys& y : sys_)
            skin = std::min(skin, y.skinNow);
        out[0] = std::sqrt(lastD2_);
        out[1] = std::sqrt(lastS2_);
        out[2] = skin;
        out[3] = (double) repeats_;
        out[4] = (double) shortSolves_;
    }
    long repeatedLaunches() const { return repeats_; }
    // The caller's stream waits on the GPU for the last evaluation's outputs (no host wait).
    void waitOn(hipStream_t stream) {
        withDevice([&] { HIPCHK(hipStreamWaitEvent(stream, evOut_, 0)); });
    }
    const std::vector<double>& forces() const { return forcesH_; }

    void dipoles(bool induced, double* out) {
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: no atoms");
        if (!haveResults_) {
            std::fill(out, out+3*n_, 0.0);
            return;
        }
        withDevice([&] {
            tmpD_.reserve(3*sizeof(double)*n_);
            launchUnsort3(st_, n_, rankD_.as<int>(), induced ? xD_.as<double>() : permD_.as<double>(),
                          tmpD_.as<double>());
            HIPCHK(hipMemcpyAsync(out, tmpD_.p, 3*sizeof(double)*n_, hipMemcpyDeviceToHost, st_));
            HIPCHK(hipStreamSynchronize(st_));
        });
    }

    void resetHistory() { histSlots_.clear(); }

    void setHistory(int nset, const double* sets) {
        if (nset < 0 || nset > 4 || (nset > 0 && !sets))
            throw std::invalid_argument("pgm_hip: set_history takes 0..4 sets");
        if (n_ <= 0)
            throw std::logic_error("pgm_hip: set_history before setAtoms");
        withDevice([&] {
            const size_t stride = 3*(size_t) nGlobal_;
            histD_.reserve(kNHist*stride*sizeof(double));
            histSlots_.clear();
            for (int k = 0; k < nset; k++) {
                HIPCHK(hipMemcpyAsync(histD_.as<double>()+k*stride, sets+k*stride, stride*sizeof(double),
                                      hipMemcpyHostToDevice, st_));
                histSlots_.push_back(k);
            }
            HIPCHK(hipStreamSynchronize(st_));
        });
    }

    void freezeHistory(bool f) { frozen_ = f; }

    void setTiming(bool on, unsigned mask = ~0u) {
        timing_ = on;
        phaseMask_ = mask;
        resetTiming();
    }
    void resetTiming() {
        for (int k = 0; k < kNumPhases; k++) {
            phaseMs_[k] = 0;
            phaseCount_[k] = 0;
        }
    }
    void timing(double* ms, long* counts) const {
        for (int k = 0; k < kNumPhases; k++) {
            ms[k] = phaseMs_[k];
            counts[k] = phaseCount_[k];
        }
    }
    const pgm_terms& lastTerms() const { return sys_[0].terms; }

private:
    // One periodic system of the engine; a batch holds several, contiguous in atom index in both orders.
    struct Sys {
        int a0 = 0, a1 = 0;
        int grid[3] = {0, 0, 0};
        double H[3][3] = {{0}}, recip[3][3] = {{0}}, volume = 0;
        double skinNow = 0, rl = 0;
        double buildBox[3][3] = {{0}};
        CellGrid cells{};
        int cellOff = 0;            // first slot of its cell counts / starts (ncell + 1 slots)
        PmeGrid pme{};
        int group = 0, slot = 0;    // transform group (grid shape) and place in it
        long long realOff = 0;      // its real grid in the stacked grids (elements)
        long long cplxOff = 0;      // its half spectrum and influence function (elements)
        int brickOff = 0;
        bool haveInfluence = false;
        double influenceBox[3][3] = {{0}};
        double q2 = 0, Qtot = 0;
        pgm_terms terms{};
        double strain[9] = {0};
    };
    // The systems that share one grid shape: their grids are stacked and transformed together.
    struct Group {
        PmeGrid g{};
        std::vector<int> members;
        long long realOff = 0, cplxOff = 0;
        long long twOff = 0, bmodOff = 0;
        int partOff = 0, blocksPerGrid = 0;
    };

    pgm_settings s_;
    hipStream_t st_ = nullptr;
    hipEvent_t evIn_ = nullptr, evOut_ = nullptr, evSync_ = nullptr, evList_ = nullptr;
    // the PME solves of an evaluation run on st2_ beside the pair passes on st_ (fork / join by events); st2_ is
    // the device's PME stream, shared by its engines (acquirePmeStream)
    hipStream_t st2_ = nullptr, pmeSt_ = nullptr;
    bool pmeOverlap_ = true;              // PME on st2_ beside the pair passes (the autotuner measures it)
    double overlapMs_[2] = {0.0, 0.0};    // the autotuner's CG product: PME after, beside the field pass
    hipEvent_t evFork_ = nullptr, evJoin_ = nullptr;
    int n_ = 0, nGlobal_ = 0, covN_ = -1, covBuiltN_ = -1;   // covBuiltN_: the atom count of the covalent CSR
    int setupLanes_ = 0;   // the setup pass's lanes when the autotuner chose them apart (0: lanes_)
    bool polarizable_ = false;
    double rgMax_ = 0.0;   // the largest radius parameter 2R^2 (orders the fused setup's segments)
    bool listReverse_ = false;   // the block order of the next pass over the compacted list
    std::vector<double> qH_, forcesH_;
    std::vector<int> gidxH_, covAtomH_, covPartnerH_, covEntryOfH_;   // covEntryOfH_: input entry -> CSR slot
    std::vector<int> giTmp_, newOfTmp_;   // setAtoms' working arrays
    bool gidxIdentity_ = false;           // gidxH_[i] == i for every i
    // systems and the PME layout
    std::vector<Sys> sys_;
    std::vector<Group> groups_;
    bool batch_ = false, layoutDirty_ = true, tablesDirty_ = true;
    long long realTotal_ = 0, cplxTotal_ = 0, twTotal_ = 0, bmodTotal_ = 0;
    int brickTotal_ = 0, cellTotal_ = 0, partTotal_ = 0;
    int brick_ = kBrickSmall;   // edge of the spreading bricks, chosen by buildLayout
    // list state
    bool haveList_ = false;
    long listBuilds_ = 0, listRenames_ = 0;
    // the asynchronous launch (launchDevice / finishDevice): its state between the two calls
    bool inFlight_ = false, asyncEval_ = false, asyncCg_ = false, syncCgNext_ = false, repeatNext_ = false;
    bool pendingChecking_ = false, pendingPrev_ = false, prevValid_ = false;
    int asyncPredict_ = 0, asyncStop_ = 0, pendingPush_ = -1;
    long shortSolves_ = 0;   // synchronous evaluations whose solve needed more than was enqueued ahead
    double lastD2_ = 0, lastS2_ = 0;   // squared: displacement since the build at the last step, the last move
    long lateChecks_ = 0, repeats_ = 0;
    int recentIts_[16] = {0}, numRecent_ = 0, recentPos_ = 0;
    hipEvent_t evDone_ = nullptr;
    DevBuf posPrevD_;
    HostBuf cgH_;
    DevBuf newOfD_, order2D_, rank2D_, build2D_;
    long long nCand_ = 0;
    int lastIterations_ = 0, iterations_ = 0;
    double residual_ = 0;
    std::deque<int> histSlots_;
    bool frozen_ = false;
    bool haveResults_ = false;
    int lanes_ = 16;
    bool fused_ = false;
    bool fieldOtf_ = false;        // FP32: the field pass evaluates its kernels instead of reading the stored ones
    void* scratchPtr_ = nullptr;   // this engine's device's static scratch (scratchAddress)
    bool noHistory_ = false;       // the allocation-free evaluation: alpha E guess, no history entry
    double tuneMs_[6] = {0, 0, 0, 0, 0, 0};
    // kernel timing
    struct Timed {
        int phase, iteration;
        hipEvent_t a, b;
    };
    bool timing_ = false;
    // GVDW (combination rule), set after setAtoms
    bool gvdwOn_ = false, gvdwSlater_ = false;
    bool gvdwReady_ = false;   // the sorted per-atom GVDW parameters are gathered (by prep, for this setting)
    double rvdw_ = 0;
    int nSkip_ = 0;
    DevBuf gvInD_, gvTD_, skipStartD_, skipPartD_, gvdwAtomD_;
    // the caller's pair list (setCandidates): per-atom segments of packed entries, caller's order
    bool extList_ = false;
    std::vector<int> extStartH_, extCandH_;
    // CG preconditioner
    double preRadius_ = 0;
    bool jacobiSelf_ = false;
    bool preReady_ = false;   // the short list of this evaluation's pairs is built (by run)
    DevBuf srCountD_, srJD_, srGD_, srOverD_;
    unsigned phaseMask_ = ~0u;   // phases bracketed with events while timing (bit k: phase k)
    std::vector<hipEvent_t> evPool_;
    size_t evUsed_ = 0;
    std::vector<Timed> timed_;
    int openPhase_ = -1, openIter_ = -1;
    hipEvent_t openEv_ = nullptr;
    double phaseMs_[kNumPhases] = {0};
    long phaseCount_[kNumPhases] = {0};

    hipEvent_t nextEvent() {
        if (evUsed_ == evPool_.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            evPool_.push_back(e);
        }
        return evPool_[evUsed_++];
    }
    // Brackets the kernels of one phase (iteration: CG iteration index, -1 outside the CG loop).
    void tb(int phase, int iteration = -1) {
        if (!timing_ || !((phaseMask_ >> phase) & 1u)) {
            openPhase_ = -1;
            return;
        }
        openPhase_ = phase;
        openIter_ = iteration;
        openEv_ = nextEvent();
        HIPCHK(hipEventRecord(openEv_, st_));
    }
    void te() {
        if (!timing_ || openPhase_ < 0)
            return;
        hipEvent_t b = nextEvent();
        HIPCHK(hipEventRecord(b, st_));
        timed_.push_back(Timed{openPhase_, openIter_, openEv_, b});
        openPhase_ = -1;
    }
    // Accumulate the launches that did work.  Each bracket's end event is waited for: most were passed by the
    // evaluation's final synchronization, but a list rebuilt by pgm_finish_device is enqueued after it.
    void collectTiming() {
        if (!timing_)
            return;
        for (const Timed& t : timed_) {
            if (t.iteration >= iterations_)
                continue;   // enqueued ahead, returned at once
            float ms = 0;
            HIPCHK(hipEventSynchronize(t.b));
            HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
            phaseMs_[t.phase] += ms;
            phaseCount_[t.phase]++;
        }
        timed_.clear();
        evUsed_ = 0;
    }
    // device buffers
    DevBuf qD_, rgD_, alphaD_, gidxD_, posInD_, posBuildD_;
    const double* posDev_ = nullptr;   // the FP64 positions (caller's order) of this evaluation
    DevBuf covStartD_, covPartnerD_, covStrengthD_, covAtomOfD_, revStartD_, revEntryD_, covInputD_;
    DevBuf cellOfD_, slotD_, cellCountD_, cellStartD_, orderD_, rankD_, sortedCellD_, scratchD_;
    DevBuf candCountD_, candStartD_, candD_, nbD_, pjD_, g12D_, g03D_, subOfD_;
    DevBuf listScanD_, brickScanD_;   // the tile sums of the long scans (list build, PME bricks)
    DevBuf qSD_, rgSD_, alphaSD_, posDD_, posFD_;
    DevBuf permD_, permTD_, edD_, erD_, bD_, xD_, rD_, zD_, pD_, apD_, tdD_, trD_, pTD_, dD_, dTD_, etotD_;
    DevBuf fpairD_, frecD_, fcovD_, forcesOutD_, tmpD_;
    DevBuf brickOfD_, pslotD_, brickCountD_, brickStartD_, pmeAtomD_, baseD_, splD_;
    DevBuf fqD_, fdD_;   // the charges and dipoles of a field evaluation (caller's order)
    DevBuf gridD_, gridFixedD_, cplxD_, inflD_, twD_, bmodD_;
    const long long* fixedClean_ = nullptr;   // gridFixedD_ as last left all zero, over fixedCleanWords_ words
    size_t fixedCleanWords_ = 0;
    DevBuf cgPartD_, partD_, resD_, cgD_, doneD_, histD_, maxBitsD_;
    DevBuf perAtomD_, chunkSumD_;
    DevBuf fftGroupsD_;     // the merged transform launches' group table
    FftLaunch fftL_{};
    // systems tables (batch)
    DevBuf sysOfD_, boxTabD_, boxTabFD_, pmeTabD_, brickSysD_, groupBoxD_, chunkStartD_, sysChunkD_;
    DevBuf recStartD_, recSysD_, qBackD_, wED_;
    int nChunks_ = 0, nRecRanges_ = 0;
    // parameter gradients
    DevBuf pairOutD_, pmeOutD_, gridTD_, gSD_, gTD_, pdotD_, lamD_, adjBD_, ddotD_, ddotTD_, etotDotD_, gInD_, gradD_;
    int adjIterations_ = 0, lastAdjIterations_ = 0;
    double adjResidual_ = 0;
    HostBuf resH_, stageH_, scalH_, listH_, cmpH_;
    HostBuf paramH_, covH_, tabH_;   // staging of setAtoms' parameters, setCovalent's strengths, the batch tables
    hipEvent_t evCov_ = nullptr, evTab_ = nullptr;   // after the last copy out of covH_ / tabH_
    hipEvent_t evParam_ = nullptr;                   // after the last copies out of paramH_
    // the last evaluation's positions and boxes, for a backward at the same positions
    DevBuf lastPosD_, cmpPosD_, cmpFlagD_;
    std::vector<double> lastBoxes_;
    bool adjointReady_ = false;   // the stored state is that evaluation's (cleared by the pair passes' timing)

    void validate() {
        const pgm_settings& s = s_;
        if (!(s.cutoff > 0))
            throw std::invalid_argument("pgm_hip: cutoff must be positive");
        if (!(s.beta > 0))
            throw std::invalid_argument("pgm_hip: Ewald coefficient must be positive");
        if (s.order < 4 || s.order > 8)
            throw std::invalid_argument("pgm_hip: PME order must be 4..8");
        for (int a = 0; a < 3; a++)
            if (s.grid[a] < s.order)
                throw std::invalid_argument("pgm_hip: PME grid must be at least the order in every dimension");
        for (int a = 0; a < 3; a++)
            (void) makePlan(s.grid[a]);
        if (!(s.tol > 0))
            throw std::invalid_argument("pgm_hip: tolerance must be positive");
        if (s.max_iter < 0)
            throw std::invalid_argument("pgm_hip: max_iter must be >= 0");
        if (s.predictor < 0 || s.predictor > 2)
            throw std::invalid_argument("pgm_hip: predictor must be 0, 1 or 2");
        if (s.skin < 0)
            throw std::invalid_argument("pgm_hip: skin must be >= 0");
        if (s.precision != PGM_DOUBLE && s.precision != PGM_MIXED)
            throw std::invalid_argument("pgm_hip: precision must be PGM_DOUBLE or PGM_MIXED");
    }

    void upload(DevBuf& b, const void* src, size_t bytes) {
        b.reserve(bytes);
        if (bytes)
            HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
    }
    // Small host -> device copies through pinned staging, without a wait: the copies are ordered on the stream
    // before every later kernel, and the staging is written again only after the event of its last copies.  The
    // allocation-free evaluation, which may not allocate pinned memory, takes the plain copies.
    struct Staged {
        DevBuf* dst;
        const void* src;
        size_t bytes;
    };
    void uploadStaged(std::initializer_list<Staged> parts) {
        if (t_arena) {
            for (const Staged& s : parts)
                upload(*s.dst, s.src, s.bytes);
            return;
        }
        size_t total = 0;
        for (const Staged& s : parts)
            total += (s.bytes+255) & ~(size_t) 255;
        HIPCHK(hipEventSynchronize(evTab_));
        tabH_.reserve(std::max<size_t>(total, 256));
        char* h = tabH_.as<char>();
        size_t off = 0;
        for (const Staged& s : parts) {
            s.dst->reserve(s.bytes);
            if (s.bytes) {
                std::memcpy(h+off, s.src, s.bytes);
                HIPCHK(hipMemcpyAsync(s.dst->p, h+off, s.bytes, hipMemcpyHostToDevice, st_));
            }
            off += (s.bytes+255) & ~(size_t) 255;
        }
        HIPCHK(hipEventRecord(evTab_, st_));
    }

    // The list kept over a renaming of the caller's indices (newOf: old -> new): order, rank and the build
    // positions renamed, the sorted parameters gathered again from the new arrays.
    // Nothing here waits for the stream.
    // the tile-sum scratch of a scan of m elements in b (null for a short scan, which runs in one block)
    int* scanScratch(DevBuf& b, int m) {
        const int k = scanScratchInts(m);
        if (!k)
            return nullptr;
        b.reserve(k*sizeof(int));
        return b.as<int>();
    }
    void renameList(const std::vector<int>& newOf) {
        const int n = n_;
        upload(newOfD_, newOf.data(), n*sizeof(int));
        order2D_.reserve(n*sizeof(int));
        rank2D_.reserve(n*sizeof(int));
        build2D_.reserve(3*n*sizeof(double));
        launchRename(st_, n, newOfD_.as<int>(), orderD_.as<int>(), rankD_.as<int>(), posBuildD_.as<double>(),
                     order2D_.as<int>(), rank2D_.as<int>(), build2D_.as<double>());
        orderD_.swap(order2D_);
        rankD_.swap(rank2D_);
        posBuildD_.swap(build2D_);
        listRenames_++;
        permuteParameters();
    }
    // the sorted copies of q, alpha and the radius parameter, gathered again from the caller-order arrays
    void permuteParameters() {
        launchPermute(st_, n_, orderD_.as<int>(), qD_.as<double>(), qSD_.as<double>());
        launchPermute(st_, n_, orderD_.as<int>(), alphaD_.as<double>(), alphaSD_.as<double>());
        launchPermute(st_, n_, orderD_.as<int>(), rgD_.as<double>(), rgSD_.as<double>());
    }

    int systemOf(int atom) const {
        for (size_t s = 0; s < sys_.size(); s++)
            if (atom >= sys_[s].a0 && atom < sys_[s].a1)
                return (int) s;
        return -1;
    }

    void refreshCharges() {
        tablesDirty_ = true;
        for (Sys& y : sys_) {
            y.q2 = y.Qtot = 0;
            for (int i = y.a0; i < y.a1 && i < (int) qH_.size(); i++) {
                y.q2 += qH_[i]*qH_[i];
                y.Qtot += qH_[i];
            }
        }
    }

    void setCovalentImpl(int ncov, const int* atom, const int* partner, const double* strength) {
        // CSR by atom (entries in input order per atom) and the reverse CSR by partner
        std::vector<int> start(n_+1, 0), rstart(n_+1, 0);
        for (int m = 0; m < ncov; m++) {
            start[atom[m]+1]++;
            rstart[partner[m]+1]++;
        }
        for (int i = 0; i < n_; i++) {
            start[i+1] += start[i];
            rstart[i+1] += rstart[i];
        }
        std::vector<int> fill(start.begin(), start.end()-1), rfill(rstart.begin(), rstart.end()-1);
        std::vector<int> part(ncov), atomOf(ncov), rev(ncov);
        std::vector<double> str(ncov);
        std::vector<int> entryOf(ncov), inputOf(ncov);
        for (int m = 0; m < ncov; m++) {
            const int k = fill[atom[m]]++;
            part[k] = partner[m];
            str[k] = strength[m];
            atomOf[k] = atom[m];
            entryOf[m] = k;
            inputOf[k] = m;
        }
        for (int m = 0; m < ncov; m++)
            rev[rfill[partner[m]]++] = entryOf[m];
        upload(covStartD_, start.data(), (n_+1)*sizeof(int));
        upload(covPartnerD_, part.data(), ncov*sizeof(int));
        upload(covStrengthD_, str.data(), ncov*sizeof(double));
        upload(covAtomOfD_, atomOf.data(), ncov*sizeof(int));
        upload(revStartD_, rstart.data(), (n_+1)*sizeof(int));
        upload(revEntryD_, rev.data(), ncov*sizeof(int));
        upload(covInputD_, inputOf.data(), ncov*sizeof(int));
        covAtomH_.assign(atom, atom+ncov);
        covPartnerH_.assign(partner, partner+ncov);
        covEntryOfH_ = entryOf;
        covN_ = ncov;
        covBuiltN_ = n_;
        haveResults_ = false;
    }

    // ------------------------------------------------------------------------------------ boxes and layout

    void setBoxes(const double* boxes) {
        if (layoutDirty_)
            buildLayout();
        bool changed = false;
        for (size_t s = 0; s < sys_.size(); s++) {
            Sys& y = sys_[s];
            const double* b = boxes+9*s;
            double H[3][3];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    H[r][c] = b[3*r+c];
            if (H[0][1] != 0 || H[0][2] != 0 || H[1][2] != 0)
                throw std::invalid_argument("pgm_hip: the box must be in reduced triclinic form "
                                            "(a = (ax,0,0), b = (bx,by,0))");
            if (!(H[0][0] > 0 && H[1][1] > 0 && H[2][2] > 0))
                throw std::invalid_argument("pgm_hip: box diagonal must be positive");
            changed = changed || std::memcmp(H, y.H, sizeof(H)) != 0;
            std::memcpy(y.H, H, sizeof(H));
            y.volume = H[0][0]*H[1][1]*H[2][2];
            // recip = H^-1 of the lower-triangular H (rows = box vectors)
            double R[3][3] = {{0}};
            R[0][0] = 1.0/H[0][0];
            R[1][1] = 1.0/H[1][1];
            R[2][2] = 1.0/H[2][2];
            R[1][0] = -H[1][0]/(H[0][0]*H[1][1]);
            R[2][1] = -H[2][1]/(H[1][1]*H[2][2]);
            R[2][0] = (H[1][0]*H[2][1]-H[1][1]*H[2][0])/(H[0][0]*H[1][1]*H[2][2]);
            std::memcpy(y.recip, R, sizeof(R));
            // plane spacings and the list radius (skin shrunk in small boxes)
            double wmin = 1e300;
            for (int a = 0; a < 3; a++)
                wmin = std::min(wmin, 1.0/std::sqrt(R[0][a]*R[0][a]+R[1][a]*R[1][a]+R[2][a]*R[2][a]));
            if (s_.cutoff >= 0.5*wmin)
                throw std::invalid_argument("pgm_hip: the cutoff must be below half the smallest box plane spacing");
            y.skinNow = std::min(s_.skin, 0.5*wmin-s_.cutoff);
            y.rl = s_.cutoff+y.skinNow;
            if (!y.haveInfluence || std::memcmp(H, y.influenceBox, sizeof(H)) != 0)
                setupInfluence(y);
        }
        if (batch_ && (changed || tablesDirty_))
            uploadBoxTables();
    }

    // Grid shapes -> transform groups; offsets of every system's grids, bricks and cells; the batch tables.
    void buildLayout() {
        groups_.clear();
        long long largeBricks = 0;
        bool gridsFit = true;   // the spread wraps a tile row into the grid once: every dimension >= the brick
        for (const Sys& y : sys_) {
            largeBricks += (long long) ((y.grid[0]+kBrickLarge-1)/kBrickLarge)*((y.grid[1]+kBrickLarge-1)/kBrickLarge)*
                           ((y.grid[2]+kBrickLarge-1)/kBrickLarge);
            for (int a = 0; a < 3; a++)
                gridsFit = gridsFit && y.grid[a] >= kBrickLarge;
        }
        const long long tileEdge = kBrickLarge+s_.order-1;
        const long long tileLarge = tileEdge*tileEdge*tileEdge;
        // the spread's LDS tile holds 8 bytes per point and grid word (one word in FP32, two in FP64: 77 KiB at
        // order 6, within the 160 KiB of LDS a gfx950 workgroup may take); higher orders keep the small brick
        const bool largeFits = tileLarge*(long long) sizeof(double) <= 64*1024;
        brick_ = (largeFits && gridsFit && largeBricks >= kMinBricks) ? kBrickLarge : kBrickSmall;
        for (size_t s = 0; s < sys_.size(); s++) {
            Sys& y = sys_[s];
            PmeGrid g{};
            for (int a = 0; a < 3; a++) {
                g.K[a] = y.grid[a];
                g.plan[a] = makePlan(g.K[a]);
                g.nb[a] = (g.K[a]+brick_-1)/brick_;
            }
            g.K3h = g.K[2]/2+1;
            g.order = s_.order;
            g.brick = brick_;
            g.nbrick = g.nb[0]*g.nb[1]*g.nb[2];
            y.pme = g;
            int gi = -1;
            for (size_t k = 0; k < groups_.size(); k++)
                if (groups_[k].g.K[0] == g.K[0] && groups_[k].g.K[1] == g.K[1] && groups_[k].g.K[2] == g.K[2])
                    gi = (int) k;
            if (gi < 0) {
                groups_.push_back(Group{});
                gi = (int) groups_.size()-1;
                groups_[gi].g = g;
            }
            y.group = gi;
            y.slot = (int) groups_[gi].members.size();
            groups_[gi].members.push_back((int) s);
            y.haveInfluence = false;
        }
        realTotal_ = cplxTotal_ = twTotal_ = bmodTotal_ = 0;
        partTotal_ = 0;
        const bool mixed = s_.precision == PGM_MIXED;
        std::vector<int> recStart(1, 0), recSys;
        for (Group& G : groups_) {
            const PmeGrid& g = G.g;
            const long long nreal = (long long) g.K[0]*g.K[1]*g.K[2], nc = (long long) g.K[0]*g.K[1]*g.K3h;
            G.realOff = realTotal_;
            G.cplxOff = cplxTotal_;
            G.twOff = twTotal_;
            G.bmodOff = bmodTotal_;
            realTotal_ += nreal*(long long) G.members.size();
            cplxTotal_ += nc*(long long) G.members.size();
            twTotal_ += g.K[0]+g.K[1]+g.K[2];
            bmodTotal_ += g.K[0]+g.K[1]+g.K[2];
            const int ng = (int) G.members.size();
            G.blocksPerGrid = mixed ? convolveBlocksPerGrid<float>(g, ng) : convolveBlocksPerGrid<double>(g, ng);
            G.partOff = partTotal_;
            partTotal_ += G.blocksPerGrid*(int) G.members.size();
            for (size_t k = 0; k < G.members.size(); k++) {
                Sys& y = sys_[G.members[k]];
                y.realOff = G.realOff+(long long) k*nreal;
                y.cplxOff = G.cplxOff+(long long) k*nc;
                recStart.push_back(G.partOff+(int) (k+1)*G.blocksPerGrid);
                recSys.push_back(G.members[k]);
            }
        }
        // the merged transform launches: every group's share of each stage
        std::vector<FftGroup> fg;
        fftL_ = FftLaunch{};
        int boxOff = 0;
        for (Group& G : groups_) {
            FftGroup f{};
            f.g = G.g;
            f.ngrid = (int) G.members.size();
            f.realOff = G.realOff;
            f.cplxOff = G.cplxOff;
            f.twOff = (int) G.twOff;
            f.partOff = G.partOff;
            f.boxOff = boxOff;
            int b[4];
            size_t l[4];
            if (mixed)
                fftGroupShape<float>(f, b, l);
            else
                fftGroupShape<double>(f, b, l);
            for (int k = 0; k < 4; k++) {
                f.blk0[k] = fftL_.blocks[k];
                fftL_.blocks[k] += b[k];
                fftL_.lds[k] = std::max(fftL_.lds[k], l[k]);
            }
            if (b[2] != G.blocksPerGrid*f.ngrid)
                throw std::logic_error("pgm_hip: transform block counts disagree");
            boxOff += f.ngrid;
            fg.push_back(f);
        }
        // z and y fused when the x-planes alone give enough blocks (many small grids) and a plane fits the LDS
        fftL_.zyFused = (fftL_.blocks[3] >= 512 && fftL_.lds[3] <= 48*1024) ? 1 : 0;
        upload(fftGroupsD_, fg.data(), fg.size()*sizeof(FftGroup));
        fftL_.groups = fftGroupsD_.as<FftGroup>();
        fftL_.ngroups = (int) fg.size();
        brickTotal_ = 0;
        for (Sys& y : sys_) {
            y.brickOff = brickTotal_;
            brickTotal_ += y.pme.nbrick;
        }
        // device arrays of the layout
        const size_t tsz = mixed ? sizeof(float) : sizeof(double), t2 = 2*tsz;
        gridD_.reserve((size_t) realTotal_*tsz);
        gridFixedD_.reserve((size_t) realTotal_*sizeof(long long)*(s_.precision == PGM_MIXED ? 1 : 2));
        cplxD_.reserve((size_t) cplxTotal_*t2);
        inflD_.reserve((size_t) cplxTotal_*tsz);
        twD_.reserve((size_t) twTotal_*t2);
        std::vector<double> bm;
        for (Group& G : groups_) {
            for (int a = 0; a < 3; a++) {
                const std::vector<double> m = bsplineModuli(G.g.K[a], G.g.order);
                bm.insert(bm.end(), m.begin(), m.end());
            }
            if (mixed)
                launchTwiddles<float>(st_, G.g, twD_.as<float2>()+G.twOff);
            else
                launchTwiddles<double>(st_, G.g, twD_.as<double2>()+G.twOff);
        }
        upload(bmodD_, bm.data(), bm.size()*sizeof(double));
        nRecRanges_ = (int) recSys.size();
        upload(recStartD_, recStart.data(), recStart.size()*sizeof(int));
        upload(recSysD_, recSys.data(), recSys.size()*sizeof(int));
        // per-system sums of per-atom rows: chunks of at most kChunkAtoms atoms inside one system
        std::vector<int> chunkStart(1, 0), sysChunk(1, 0);
        for (Sys& y : sys_) {
            for (int a = y.a0; a < y.a1; a += kChunkAtoms)
                chunkStart.push_back(std::min(y.a1, a+kChunkAtoms));
            sysChunk.push_back((int) chunkStart.size()-1);
        }
        nChunks_ = (int) chunkStart.size()-1;
        upload(chunkStartD_, chunkStart.data(), chunkStart.size()*sizeof(int));
        upload(sysChunkD_, sysChunk.data(), sysChunk.size()*sizeof(int));
        if (batch_) {
            std::vector<int> sysOf(n_), brickSys(brickTotal_);
            for (size_t s = 0; s < sys_.size(); s++) {
                for (int i = sys_[s].a0; i < sys_[s].a1; i++)
                    sysOf[i] = (int) s;
                for (int b = 0; b < sys_[s].pme.nbrick; b++)
                    brickSys[sys_[s].brickOff+b] = (int) s;
            }
            upload(sysOfD_, sysOf.data(), n_*sizeof(int));
            upload(brickSysD_, brickSys.data(), brickTotal_*sizeof(int));
        }
        layoutDirty_ = false;
        tablesDirty_ = true;
        haveList_ = false;
    }

    void setupInfluence(Sys& y) {
        const Group& G = groups_[y.group];
        const BoxT<double> bd = makeBox<double>(y.H, y.recip);
        if (s_.precision == PGM_MIXED)
            launchInfluence<float>(st_, y.pme, bmodD_.as<double>()+G.bmodOff, bd, s_.beta, y.volume,
                                   inflD_.as<float>()+y.cplxOff);
        else
            launchInfluence<double>(st_, y.pme, bmodD_.as<double>()+G.bmodOff, bd, s_.beta, y.volume,
                                    inflD_.as<double>()+y.cplxOff);
        std::memcpy(y.influenceBox, y.H, sizeof(y.H));
        y.haveInfluence = true;
    }

    void uploadBoxTables() {
        const size_t S = sys_.size();
        std::vector<BoxT<double>> bd(S), gb;
        std::vector<BoxT<float>> bf(S);
        std::vector<PmeSys> ps(S);
        std::vector<double> qb(S);
        for (size_t s = 0; s < S; s++) {
            const Sys& y = sys_[s];
            bd[s] = makeBox<double>(y.H, y.recip);
            bf[s] = makeBox<float>(y.H, y.recip);
            ps[s].g = y.pme;
            ps[s].box = bd[s];
            ps[s].gridOff = y.realOff;
            ps[s].brickOff = y.brickOff;
            ps[s].pad = 0;
            qb[s] = kPi*y.Qtot/(y.volume*s_.beta*s_.beta);
        }
        for (const Group& G : groups_)
            for (int m : G.members)
                gb.push_back(bd[m]);
        uploadStaged({{&boxTabD_, bd.data(), S*sizeof(BoxT<double>)},
                      {&boxTabFD_, bf.data(), S*sizeof(BoxT<float>)},
                      {&pmeTabD_, ps.data(), S*sizeof(PmeSys)},
                      {&groupBoxD_, gb.data(), gb.size()*sizeof(BoxT<double>)},
                      {&qBackD_, qb.data(), S*sizeof(double)}});
        tablesDirty_ = false;
    }

    template<class T>
    BoxSel<T> boxSel() const {
        BoxSel<T> b{};
        b.ext = extList_ ? 1 : 0;
        if (batch_) {
            b.table = sizeof(T) == sizeof(float) ? reinterpret_cast<const BoxT<T>*>(boxTabFD_.p)
                                                 : reinterpret_cast<const BoxT<T>*>(boxTabD_.p);
            b.sysOf = sysOfD_.as<int>();
        }
        else {
            b.box = makeBox<T>(sys_[0].H, sys_[0].recip);
            b.table = nullptr;
            b.sysOf = nullptr;
        }
        return b;
    }

    PmeSel pmeSel() const {
        PmeSel p{};
        p.g = sys_[0].pme;
        p.box = makeBox<double>(sys_[0].H, sys_[0].recip);
        p.table = batch_ ? pmeTabD_.as<PmeSys>() : nullptr;
        p.sysOf = batch_ ? sysOfD_.as<int>() : nullptr;
        p.brickSys = batch_ ? brickSysD_.as<int>() : nullptr;
        p.nbrick = brickTotal_;
        p.order = s_.order;
        p.brick = brick_;
        p.fixedStride = (size_t) realTotal_;
        return p;
    }

    void stagePositions(PosSource src, const void* pos) {
        const size_t bytes = 3*sizeof(double)*n_;
        switch (src) {
        case PosSource::HostF64:
            posInD_.reserve(bytes);
            stageH_.reserve(bytes);
            std::memcpy(stageH_.p, pos, bytes);
            HIPCHK(hipMemcpyAsync(posInD_.p, stageH_.p, bytes, hipMemcpyHostToDevice, st_));
            posDev_ = posInD_.as<double>();
            break;
        case PosSource::HostF32: {
            posInD_.reserve(bytes);
            stageH_.reserve(bytes);
            const float* f = static_cast<const float*>(pos);
            double* d = stageH_.as<double>();
            for (int k = 0; k < 3*n_; k++)
                d[k] = f[k];
            HIPCHK(hipMemcpyAsync(posInD_.p, stageH_.p, bytes, hipMemcpyHostToDevice, st_));
            posDev_ = posInD_.as<double>();
            break;
        }
        case PosSource::DeviceF64:
            posDev_ = static_cast<const double*>(pos);
            break;
        case PosSource::DeviceF4:
            posInD_.reserve(bytes);
            launchF4ToDouble(st_, n_, static_cast<const float4*>(pos), posInD_.as<double>());
            posDev_ = posInD_.as<double>();
            break;
        }
    }

    // ------------------------------------------------------------------------------------ lists

    // Host wait for the work enqueued so far on the engine stream, spinning on an event: a blocking
    // synchronize adds an interrupt's wake-up latency to each of an evaluation's few round trips.
    void waitEvent(hipEvent_t ev) {
        hipError_t e;
        while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
        }
        HIPCHK(e);
    }
    void waitStream() {
        HIPCHK(hipEventRecord(evSync_, st_));
        waitEvent(evSync_);
    }

    // The list check in two halves: the largest displacement since the build is copied to pinned host memory
    // without waiting (the maximum starts at zero: every preparation clears it for the next check; a value
    // left over by an interrupted evaluation can only be larger, so at worst it rebuilds early).
    // withPrev: also the largest move since the previous launch's positions (posPrevD_), into the second word.
    void startListCheck(bool withPrev = false) {
        launchMaxDisplacement(st_, n_, posDev_, posBuildD_.as<double>(), boxSel<double>(), maxBitsD_.as<unsigned>());
        if (withPrev)
            launchMaxDisplacement(st_, n_, posDev_, posPrevD_.as<double>(), boxSel<double>(),
                                  maxBitsD_.as<unsigned>()+1);
        listH_.reserve(64);
        HIPCHK(hipMemcpyAsync(listH_.p, maxBitsD_.p, 2*sizeof(unsigned), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipEventRecord(evList_, st_));
    }
    bool listStale() {
        waitEvent(evList_);
        float d2;
        std::memcpy(&d2, listH_.p, sizeof(float));
        return staleFor(d2, 0.0);
    }
    // A list built for displacements below skin/2 is stale when twice the largest displacement since the build
    // (plus ahead, a predicted further move) and the box change exceed the skin (core/cuda/pgm_cuda.cu:1898-1928).
    bool staleFor(double d2, double ahead) const {
        double boxMove = 0, skin = 1e300;
        for (const Sys& y : sys_) {
            double dH = 0;
            for (int a = 0; a < 3; a++) {
                double r2 = 0;
                for (int c = 0; c < 3; c++)
                    r2 += (y.H[a][c]-y.buildBox[a][c])*(y.H[a][c]-y.buildBox[a][c]);
                dH += std::sqrt(r2);
            }
            boxMove = std::max(boxMove, dH);
            skin = std::min(skin, y.skinNow);
        }
        return 2.0*(std::sqrt(d2)+ahead)+boxMove > skin;
    }

    // The caller's pair list in place of the engine's own: the caller's order kept (identity permutation), its
    // per-atom segments uploaded as the candidates; no displacement checks follow.
    void buildExternalList() {
        const int n = n_;
        std::vector<int> ident(n);
        std::iota(ident.begin(), ident.end(), 0);
        upload(orderD_, ident.data(), n*sizeof(int));
        upload(rankD_, ident.data(), n*sizeof(int));
        qSD_.reserve(n*sizeof(double));
        alphaSD_.reserve(n*sizeof(double));
        rgSD_.reserve(n*sizeof(double));
        launchPermute(st_, n, orderD_.as<int>(), qD_.as<double>(), qSD_.as<double>());
        launchPermute(st_, n, orderD_.as<int>(), alphaD_.as<double>(), alphaSD_.as<double>());
        launchPermute(st_, n, orderD_.as<int>(), rgD_.as<double>(), rgSD_.as<double>());
        posDD_.reserve(n*sizeof(double4));
        launchGatherPositions(st_, n, orderD_.as<int>(), posDev_, qSD_.as<double>(), boxSel<double>(),
                              posDD_.as<double4>(), nullptr);
        const int total = extStartH_[n];
        nCand_ = total;
        upload(candStartD_, extStartH_.data(), (n+1)*sizeof(int));
        const size_t cap = (size_t) total+64;
        const size_t g2 = (s_.precision == PGM_MIXED ? sizeof(float2) : sizeof(double2));
        candD_.reserve(cap*sizeof(int));
        HIPCHK(hipMemcpyAsync(candD_.p, extCandH_.data(), std::max(total, 1)*sizeof(int), hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
        pjD_.reserve(cap*sizeof(int));
        g12D_.reserve(cap*g2);
        if (s_.precision == PGM_DOUBLE)
            g03D_.reserve(cap*g2);
        for (Sys& y : sys_)
            std::memcpy(y.buildBox, y.H, sizeof(y.H));
        haveList_ = true;
        listBuilds_++;
    }

    void buildList() {
        if (extList_) {
            buildExternalList();
            return;
        }
        const int n = n_;
        // cells at least rl/2 wide, +-2 stencil (all distinct cells when fewer than 5 per dimension), per system
        cellTotal_ = 0;
        for (Sys& y : sys_) {
            CellGrid g{};
            for (int a = 0; a < 3; a++) {
                const double w =
                    1.0/std::sqrt(y.recip[0][a]*y.recip[0][a]+y.recip[1][a]*y.recip[1][a]+y.recip[2][a]*y.recip[2][a]);
                g.nc[a] = std::max(1, (int) std::floor(w/(0.5*y.rl)));
                if (g.nc[a] >= 5) {
                    g.noff[a] = 5;
                    for (int o = 0; o < 5; o++)
                        g.off[a][o] = o-2;
                }
                else {
                    g.noff[a] = g.nc[a];
                    for (int o = 0; o < g.nc[a]; o++)
                        g.off[a][o] = o;
                }
            }
            g.ncell = g.nc[0]*g.nc[1]*g.nc[2];
            y.cells = g;
            y.cellOff = cellTotal_;
            cellTotal_ += g.ncell+1;
        }
        cellOfD_.reserve(n*sizeof(int));
        slotD_.reserve(n*sizeof(int));
        cellCountD_.reserve(cellTotal_*sizeof(int));
        cellStartD_.reserve(cellTotal_*sizeof(int));
        orderD_.reserve(n*sizeof(int));
        rankD_.reserve(n*sizeof(int));
        sortedCellD_.reserve(n*sizeof(int));
        subOfD_.reserve(n*sizeof(int));
        launchZero<int>(st_, cellCountD_.as<int>(), (size_t) cellTotal_, nullptr);
        for (const Sys& y : sys_) {
            const int ns = y.a1-y.a0;
            const BoxT<double> bd = makeBox<double>(y.H, y.recip);
            int* cellCount = cellCountD_.as<int>()+y.cellOff;
            int* cellStart = cellStartD_.as<int>()+y.cellOff;
            launchCellAssign(st_, ns, posDev_+3*(size_t) y.a0, bd, y.cells, cellOfD_.as<int>()+y.a0,
                             slotD_.as<int>()+y.a0, cellCount, subOfD_.as<int>()+y.a0);
            launchScan(st_, cellCount, cellStart, y.cells.ncell, scanScratch(listScanD_, y.cells.ncell));
            launchCellScatter(st_, ns, cellOfD_.as<int>()+y.a0, slotD_.as<int>()+y.a0, cellStart,
                              orderD_.as<int>()+y.a0, y.a0);
            launchCellSort(st_, y.cells.ncell, cellStart, subOfD_.as<int>(), orderD_.as<int>()+y.a0);
            launchRank(st_, ns, orderD_.as<int>()+y.a0, rankD_.as<int>(), sortedCellD_.as<int>()+y.a0,
                       cellOfD_.as<int>(), y.a0);
        }
        // sorted parameters
        qSD_.reserve(n*sizeof(double));
        alphaSD_.reserve(n*sizeof(double));
        rgSD_.reserve(n*sizeof(double));
        tmpD_.reserve(3*n*sizeof(double));
        launchPermute(st_, n, orderD_.as<int>(), qD_.as<double>(), qSD_.as<double>());
        launchPermute(st_, n, orderD_.as<int>(), alphaD_.as<double>(), alphaSD_.as<double>());
        launchPermute(st_, n, orderD_.as<int>(), rgD_.as<double>(), rgSD_.as<double>());
        posDD_.reserve(n*sizeof(double4));
        launchGatherPositions(st_, n, orderD_.as<int>(), posDev_, qSD_.as<double>(), boxSel<double>(),
                              posDD_.as<double4>(), nullptr);
        // candidates within rl, full lists in per-atom segments
        candCountD_.reserve(n*sizeof(int));
        candStartD_.reserve((n+1)*sizeof(int));
        for (const Sys& y : sys_) {
            const BoxT<double> bd = makeBox<double>(y.H, y.recip);
            launchCandCount(st_, y.a1-y.a0, posDD_.as<double4>()+y.a0, sortedCellD_.as<int>()+y.a0,
                            cellStartD_.as<int>()+y.cellOff, y.cells, bd, y.rl*y.rl, candCountD_.as<int>()+y.a0);
        }
        launchScan(st_, candCountD_.as<int>(), candStartD_.as<int>(), n, scanScratch(listScanD_, n));
        scalH_.reserve(64);
        HIPCHK(hipMemcpyAsync(scalH_.p, candStartD_.as<int>()+n, sizeof(int), hipMemcpyDeviceToHost, st_));
        waitStream();   // a spin on an event: no interrupt wake-up in this round trip
        int total;
        std::memcpy(&total, scalH_.p, sizeof(int));
        if (total < 0)
            throw std::runtime_error("pgm_hip: pair list larger than 2^31 entries");
        nCand_ = total;
        const size_t cap = (size_t) total+64;
        const size_t g2 = (s_.precision == PGM_MIXED ? sizeof(float2) : sizeof(double2));
        candD_.reserve(cap*sizeof(int));
        pjD_.reserve(cap*sizeof(int));
        g12D_.reserve(cap*g2);
        if (s_.precision == PGM_DOUBLE)   // the FP32 path rebuilds G0, G3 from G1, G2
            g03D_.reserve(cap*g2);
        for (const Sys& y : sys_) {
            const BoxT<double> bd = makeBox<double>(y.H, y.recip);
            launchCandFill(st_, y.a1-y.a0, posDD_.as<double4>()+y.a0, sortedCellD_.as<int>()+y.a0,
                           cellStartD_.as<int>()+y.cellOff, y.cells, bd, y.rl*y.rl, candStartD_.as<int>()+y.a0,
                           candD_.as<int>(), y.a0);
        }
        posBuildD_.reserve(3*n*sizeof(double));
        HIPCHK(hipMemcpyAsync(posBuildD_.p, posDev_, 3*n*sizeof(double), hipMemcpyDeviceToDevice, st_));
        for (Sys& y : sys_)
            std::memcpy(y.buildBox, y.H, sizeof(y.H));
        haveList_ = true;
        listBuilds_++;
    }

    GuessCoef guessCoefficients(bool newStep) const {
        GuessCoef gc{};
        if (noHistory_)
            return gc;
        const int h = (int) histSlots_.size();
        auto set = [&](std::initializer_list<double> c) {
            gc.nc = (int) c.size();
            int k = 0;
            for (double v : c) {
                gc.c[k] = v;
                gc.slot[k] = histSlots_[k];
                k++;
            }
        };
        if (!newStep && h > 0)
            set({1.0});
        else if (s_.predictor == 2 && h >= 4)
            set({4.0, -6.0, 4.0, -1.0});
        else if (s_.predictor >= 1 && h >= 3)
            set({3.0, -3.0, 1.0});
        else if (s_.predictor >= 1 && h >= 1)
            set({1.0});
        return gc;
    }

    int freeSlot() const {
        for (int k = 0; k < kNHist; k++)
            if (std::find(histSlots_.begin(), histSlots_.end(), k) == histSlots_.end())
                return k;
        return histSlots_.back();
    }

    // ------------------------------------------------------------------------------------ passes

    // Spread + transforms (per grid shape) into grid (the stacked real grids): the convolved potential grids.
    // energyMode: per-system k-space energy (and strain) into res[s][kResRec..].
    // PME beside the pair passes unless PME phases are being timed (their brackets are on the engine stream).
    bool overlapPme() const {
        constexpr unsigned pmeMask = (1u << kPhSpread) | (1u << kPhFft) | (1u << kPhGather);
        return pmeOverlap_ && !(timing_ && (phaseMask_ & pmeMask));
    }
    // The PME stream of one evaluation, back to the engine stream when the scope ends (also on an error).
    struct PmeStreamScope {
        DeviceEngine& e;
        PmeStreamScope(DeviceEngine& eng, bool overlap) : e(eng) { e.pmeSt_ = overlap ? e.st2_ : e.st_; }
        ~PmeStreamScope() { e.pmeSt_ = e.st_; }
    };
    // st2_ starts after everything enqueued on st_ so far; st_ continues after everything enqueued on st2_.
    void fork() {
        if (pmeSt_ == st_)
            return;
        HIPCHK(hipEventRecord(evFork_, st_));
        HIPCHK(hipStreamWaitEvent(st2_, evFork_, 0));
    }
    void join() {
        if (pmeSt_ == st_)
            return;
        HIPCHK(hipEventRecord(evJoin_, st2_));
        HIPCHK(hipStreamWaitEvent(st_, evJoin_, 0));
    }

    template<class T>
    void pmeGrids(const typename Vec4<T>::type* dip, bool withQ, const typename Vec4<T>::type* posq, int energyMode,
                  bool strain, T* grid, const int* skip, double* res, int it,
                  const typename Vec4<T>::type* dipDot = nullptr, const typename Vec4<T>::type* disp = nullptr) {
        typedef typename Vec2<T>::type T2;
        const PmeSel ps = pmeSel();
        // the charge/dipole grids through the fixed-point accumulator (order-independent), a tangent density
        // (dipDot) straight into grid
        long long* fixed = dipDot ? nullptr : gridFixedD_.as<long long>();
        const size_t fixedWords = (size_t) realTotal_*GridFixed<T>::words;
        tb(kPhSpread, it);
        if (fixed) {
            // the forward z transform zeroes every word it reads, so the grid is clean for the next spread; it is
            // cleared here only when that is not known (new buffer, a larger layout, a solve cut short)
            if (fixed != fixedClean_ || fixedWords > fixedCleanWords_)
                launchZero<long long>(pmeSt_, fixed, fixedWords, nullptr);
            fixedClean_ = nullptr;
        }
        else
            launchZero<T>(pmeSt_, grid, (size_t) realTotal_, skip);
        launchSpread<T>(pmeSt_, ps, brickStartD_.as<int>(), pmeAtomD_.as<int>(), baseD_.as<int>(), splD_.as<T>(), dip,
                        posq, withQ ? 1 : 0, grid, fixed, skip, dipDot, disp);
        te();
        tb(kPhFft, it);
        const BoxT<double> one = makeBox<double>(sys_[0].H, sys_[0].recip);
        launchPmeConvolve<T>(pmeSt_, fftL_, grid, cplxD_.as<T2>(), inflD_.as<T>(), twD_.as<T2>(), one,
                             batch_ ? groupBoxD_.as<BoxT<double>>() : nullptr, s_.beta, energyMode, strain ? 1 : 0,
                             partD_.as<double>(), skip, fixed, (size_t) realTotal_);
        if (fixed) {   // enqueued: the spread's words are read and zeroed before anything else uses the grid
            fixedClean_ = fixed;
            fixedCleanWords_ = fixedWords;
        }
        if (energyMode)
            launchRangeSums(pmeSt_, partD_.as<double>(), 10, recStartD_.as<int>(), nRecRanges_, recSysD_.as<int>(),
                            kResCount, res+kResRec);
        te();
    }

    template<class T>
    void pmeSolve(const typename Vec4<T>::type* dip, bool withQ, const typename Vec4<T>::type* posq, int energyMode,
                  bool strain, double* field, double* force, const int* skip, double* res, int it = -1) {
        pmeGrids<T>(dip, withQ, posq, energyMode, strain, gridD_.as<T>(), skip, res, it);
        tb(kPhGather, it);
        launchGather<T>(pmeSt_, pmeSel(), n_, pmeAtomD_.as<int>(), baseD_.as<int>(), splD_.as<T>(), gridD_.as<T>(), dip,
                        posq, force ? 1 : 0, field, force, skip);
        te();
    }

    template<class T>
    void matvec(const typename Vec4<T>::type* posT, const typename Vec4<T>::type* dipT, const int* skip,
                int it = -1) {
        // the fork is recorded before the pair pass is enqueued (the PME stream waits for nothing later), and the
        // long pair pass is enqueued first, so the host's PME launches do not delay it
        fork();
        tb(kPhPairField, it);
        launchPairField<T>(st_, n_, posT, dipT, candStartD_.as<int>(), nbD_.as<int>(), pjD_.as<int>(),
                           g12D_.as<typename Vec2<T>::type>(), boxSel<T>(), tdD_.as<double>(), skip, lanes_, fieldOtf_,
                           (T) s_.beta, nextListDirection());
        te();
        pmeSolve<T>(dipT, false, posT, 0, false, trD_.as<double>(), nullptr, skip, nullptr, it);
        join();
    }

    // Jacobi-preconditioned CG on the polarizable rows of (alpha^-1 - T) x = b from the guess x, whose product
    // T x is in tdD_ / trD_.  Iterations are enqueued ahead in chunks of the last solve's count, each behind a
    // device flag, and the host reads the state once per chunk.  Returns the iterations; err: the final residual.
    template<class T>
    int cgSolve(const typename Vec4<T>::type* posT, const double* b, double* x, int& lastIts, double& err) {
        typedef typename Vec4<T>::type T4;
        typedef typename Vec2<T>::type T2;
        const int n = n_;
        const double cself = 4.0*s_.beta*s_.beta*s_.beta/(3.0*kSqrtPi);
        const Precond pc{preReady_, (!preReady_ && jacobiSelf_) ? cself : 0.0};
        const BoxSel<T> bt = boxSel<T>();
        auto precond = [&](int it, bool init) {
            if (pc.pre)
                launchPrecond<T>(st_, n, it, init, alphaSD_.as<double>(), rD_.as<double>(), apD_.as<double>(), posT,
                                 bt, srCountD_.as<int>(), srJD_.as<int>(), srGD_.as<T2>(), srOverD_.as<int>(),
                                 zD_.as<double>(), pD_.as<double>(), pTD_.as<T4>(), cgPartD_.as<double>(),
                                 doneD_.as<int>(), rgT<T>());
        };
        tb(kPhCg);
        launchCgInit<T>(st_, n, alphaSD_.as<double>(), b, x, tdD_.as<double>(), trD_.as<double>(), cself, s_.tol,
                        s_.max_iter, rD_.as<double>(), zD_.as<double>(), pD_.as<double>(), pTD_.as<T4>(),
                        cgPartD_.as<double>(), cgD_.as<CgState>(), doneD_.as<int>(), pc, rgT<T>());
        precond(0, true);
        te();
        int it = 0;
        if (asyncCg_) {
            // enqueued ahead, no host wait: the predicted iterations, then the state and the flag that says
            // whether they sufficed, read by the launch's completion
            const int stop = std::min(s_.max_iter, asyncPredict_);
            for (; it < stop; it++)
                cgIteration<T>(posT, x, it, pc);
            cgH_.reserve(64);
            HIPCHK(hipMemcpyAsync(cgH_.p, cgD_.p, sizeof(CgState), hipMemcpyDeviceToHost, st_));
            HIPCHK(hipMemcpyAsync(cgH_.as<char>()+sizeof(CgState), doneD_.as<int>()+stop, sizeof(int),
                                  hipMemcpyDeviceToHost, st_));
            asyncStop_ = stop;
            return -1;
        }
        int chunk = std::max(1, lastIts);
        scalH_.reserve(64);
        while (true) {
            const int stop = std::min(s_.max_iter, it+chunk);
            for (; it < stop; it++)
                cgIteration<T>(posT, x, it, pc);
            HIPCHK(hipMemcpyAsync(scalH_.p, cgD_.p, sizeof(CgState), hipMemcpyDeviceToHost, st_));
            HIPCHK(hipMemcpyAsync(scalH_.as<char>()+sizeof(CgState), doneD_.as<int>()+it, sizeof(int),
                                  hipMemcpyDeviceToHost, st_));
            waitStream();
            int doneNow;
            std::memcpy(&doneNow, scalH_.as<char>()+sizeof(CgState), sizeof(int));
            if (doneNow || it >= s_.max_iter)
                break;
            chunk = 1;
        }
        CgState cs;
        std::memcpy(&cs, scalH_.p, sizeof(CgState));
        lastIts = cs.iterations;
        err = cs.err;
        noteIterations(cs.iterations);
        return cs.iterations;
    }

    // One CG iteration behind the device flag done[it] (every kernel returns at once when it is set).
    template<class T>
    void cgIteration(const typename Vec4<T>::type* posT, double* x, int it, const Precond& pc) {
        typedef typename Vec4<T>::type T4;
        typedef typename Vec2<T>::type T2;
        const int n = n_;
        const double cself = 4.0*s_.beta*s_.beta*s_.beta/(3.0*kSqrtPi);
        const int* skip = doneD_.as<int>()+it;
        matvec<T>(posT, pTD_.as<T4>(), skip, it);
        tb(kPhCg, it);
        launchCgAp(st_, n, it, alphaSD_.as<double>(), pD_.as<double>(), tdD_.as<double>(), trD_.as<double>(), cself,
                   apD_.as<double>(), cgPartD_.as<double>(), doneD_.as<int>());
        launchCgUpdate(st_, n, it, alphaSD_.as<double>(), pD_.as<double>(), apD_.as<double>(), x, rD_.as<double>(),
                       zD_.as<double>(), cgPartD_.as<double>(), doneD_.as<int>(), pc);
        if (pc.pre)
            launchPrecond<T>(st_, n, it, false, alphaSD_.as<double>(), rD_.as<double>(), apD_.as<double>(), posT,
                             boxSel<T>(), srCountD_.as<int>(), srJD_.as<int>(), srGD_.as<T2>(), srOverD_.as<int>(),
                             zD_.as<double>(), pD_.as<double>(), pTD_.as<T4>(), cgPartD_.as<double>(),
                             doneD_.as<int>(), rgT<T>());
        launchCgDirection<T>(st_, n, it, s_.tol, s_.max_iter, zD_.as<double>(), pD_.as<double>(), pTD_.as<T4>(),
                             cgPartD_.as<double>(), cgD_.as<CgState>(), doneD_.as<int>(), rgT<T>());
        te();
    }

    // The iterations the last 16 solves needed: the CG enqueued ahead takes their maximum (one more if they
    // varied).
    void noteIterations(int need) {
        recentIts_[recentPos_] = need;
        recentPos_ = (recentPos_+1)%16;
        numRecent_ = std::min(numRecent_+1, 16);
    }
    // That maximum, and one more when the counts varied: a solve that needs more than were enqueued repeats the
    // whole evaluation, while an iteration enqueued but not needed still dispatches every workgroup of its
    // kernels (each returns at once), about 40 us at 96k atoms, which a steady count need not pay.
    int predictedIterations() const {
        int hi = 1, lo = 1 << 30;
        for (int k = 0; k < numRecent_; k++) {
            hi = std::max(hi, recentIts_[k]);
            lo = std::min(lo, recentIts_[k]);
        }
        return std::min(hi+(numRecent_ > 0 && lo < hi ? 1 : 0), std::max(s_.max_iter, 1));
    }

    // Per-system sums of per-atom rows [n][width] into res[s][slot..slot+width).
    void systemSums(const double* rows, int width, double* res, int slot) {
        chunkSumD_.reserve((size_t) std::max(nChunks_, 1)*11*sizeof(double));
        launchRangeSums(st_, rows, width, chunkStartD_.as<int>(), nChunks_, nullptr, width, chunkSumD_.as<double>());
        launchRangeSums(st_, chunkSumD_.as<double>(), width, sysChunkD_.as<int>(), (int) sys_.size(), nullptr,
                        kResCount, res+slot);
    }

    template<class T>
    typename Vec4<T>::type* posT() {
        return reinterpret_cast<typename Vec4<T>::type*>(sizeof(T) == sizeof(float) ? posFD_.p : posDD_.p);
    }
    template<class T>
    const T* rgT() {
        return sizeof(T) == sizeof(float) ? scratchD_.as<T>() : rgSD_.as<T>();
    }

    // The block order of the next pass over the compacted list, alternated pass by pass: each pass starts on the
    // part of the list (and of the partner data) that the pass before reached last.
    bool nextListDirection() {
        const bool r = listReverse_;
        listReverse_ = !r;
        return r;
    }

    // The candidates within rc, their kernels and the field of (q, perm) (and of dual): one fused pass or
    // compaction + setup.
    template<class T>
    void pairSetup(int lanes, bool fused, const typename Vec4<T>::type* dual) {
        typedef typename Vec4<T>::type T4;
        typedef typename Vec2<T>::type T2;
        const BoxSel<T> bt = boxSel<T>();
        const T rc2 = (T) (s_.cutoff*s_.cutoff);
        listReverse_ = true;   // the setup writes the list forward: the next pass reads it from the end
        if (fused) {
            tb(kPhPairSetup);
            launchPairCompactSetup<T>(st_, n_, posT<T>(), rgT<T>(), permTD_.as<T4>(), dual, candStartD_.as<int>(),
                                      candD_.as<int>(), nbD_.as<int>(), pjD_.as<int>(), g12D_.as<T2>(),
                                      g03D_.as<T2>(), bt, rc2, (T) s_.beta, (T) rgMax_, edD_.as<double>(),
                                      tdD_.as<double>(), lanes);
            te();
            return;
        }
        tb(kPhPairCompact);
        launchPairCompact<T>(st_, n_, posT<T>(), candStartD_.as<int>(), candD_.as<int>(), nbD_.as<int>(),
                             pjD_.as<int>(), bt, rc2, lanes);
        te();
        tb(kPhPairSetup);
        launchPairSetup<T>(st_, n_, posT<T>(), rgT<T>(), permTD_.as<T4>(), dual, candStartD_.as<int>(),
                           nbD_.as<int>(), pjD_.as<int>(), g12D_.as<T2>(), g03D_.as<T2>(), bt, (T) s_.beta,
                           edD_.as<double>(), tdD_.as<double>(), lanes);
        te();
    }

    template<class T>
    GvdwArgs<T> gvdwArgs() {
        GvdwArgs<T> g{};
        if (!gvdwOn_ || !gvdwReady_)
            return g;
        g.gv = gvTD_.as<typename Vec4<T>::type>();
        g.rvdw2 = (T) (rvdw_*rvdw_);
        g.slater = gvdwSlater_ ? 1 : 0;
        g.invKe = (T) (1.0/s_.ke);
        g.energy = gvdwAtomD_.as<double>();
        return g;
    }

    // Per-evaluation preparation in the current sorted order: positions, charges, radii and permanent dipoles
    // of the sorted atoms (one fused kernel that also clears the PME brick counters, the result sums and the
    // next list check's maximum), then the PME bricks and splines.
    template<class T>
    void prep() {
        typedef typename Vec4<T>::type T4;
        const int n = n_;
        for (DevBuf* b : {&permD_, &edD_})
            b->reserve(3*sizeof(double)*n);
        permTD_.reserve(n*sizeof(T4));
        posDD_.reserve(n*sizeof(double4));
        const bool f32 = sizeof(T) == sizeof(float);
        if (f32) {
            posFD_.reserve(n*sizeof(float4));
            scratchD_.reserve(n*sizeof(float));
        }
        brickOfD_.reserve(n*sizeof(int));
        pslotD_.reserve(n*sizeof(int));
        brickCountD_.reserve((brickTotal_+1)*sizeof(int));
        brickStartD_.reserve((brickTotal_+1)*sizeof(int));
        pmeAtomD_.reserve(n*sizeof(int));
        baseD_.reserve(3*n*sizeof(int));
        splD_.reserve((size_t) n*9*s_.order*sizeof(T));
        resD_.reserve(sys_.size()*kResCount*sizeof(double));
        maxBitsD_.reserve(2*sizeof(unsigned));
        tb(kPhPrep);
        PrepArgs a{};
        a.n = n;
        a.order = orderD_.as<int>();
        a.pos = posDev_;
        a.qS = qSD_.as<double>();
        a.rgS = rgSD_.as<double>();
        a.covStart = covStartD_.as<int>();
        a.covPartner = covPartnerD_.as<int>();
        a.covStrength = covStrengthD_.as<double>();
        a.box = boxSel<double>();
        a.posD = posDD_.as<double4>();
        a.posF = f32 ? posFD_.as<float4>() : nullptr;
        a.rgF = f32 ? scratchD_.as<float>() : nullptr;
        a.perm = permD_.as<double>();
        a.permT = permTD_.p;
        a.zeroI = brickCountD_.as<int>();
        a.nZeroI = brickTotal_+1;
        a.zeroD = resD_.as<double>();
        a.nZeroD = (int) (sys_.size()*kResCount);
        a.maxBits = maxBitsD_.as<unsigned>();
        if (gvdwOn_) {
            gvTD_.reserve(n*sizeof(T4));
            gvdwAtomD_.reserve(n*sizeof(double));
            a.gvIn = gvInD_.as<double>();
            a.gvOut = gvTD_.p;
        }
        launchPrepAtoms<T>(st_, a);
        gvdwReady_ = gvdwOn_;
        const PmeSel ps = pmeSel();
        launchPmeBin<T>(st_, n, posDD_.as<double4>(), ps, brickOfD_.as<int>(), pslotD_.as<int>(),
                        brickCountD_.as<int>());
        launchScan(st_, brickCountD_.as<int>(), brickStartD_.as<int>(), brickTotal_,
                   scanScratch(brickScanD_, brickTotal_));
        launchPmePlace<T>(st_, n, posDD_.as<double4>(), ps, brickOfD_.as<int>(), pslotD_.as<int>(),
                          brickStartD_.as<int>(), pmeAtomD_.as<int>(), baseD_.as<int>(), splD_.as<T>());
        te();
    }

    template<class T>
    void run(bool newStep, bool wantStrain) {
        typedef typename Vec4<T>::type T4;
        typedef typename Vec2<T>::type T2;
        const int n = n_;
        const size_t v3 = 3*sizeof(double)*n;
        const BoxSel<double> bd = boxSel<double>();
        const BoxSel<T> bt = boxSel<T>();
        const double cself = 4.0*s_.beta*s_.beta*s_.beta/(3.0*kSqrtPi);
        // per-atom buffers
        for (DevBuf* b : {&permD_, &edD_, &erD_, &bD_, &xD_, &rD_, &zD_, &pD_, &apD_, &tdD_, &trD_, &dD_, &etotD_,
                          &fpairD_, &frecD_, &fcovD_})
            b->reserve(v3);
        for (DevBuf* b : {&permTD_, &pTD_, &dTD_})
            b->reserve(n*sizeof(T4));
        perAtomD_.reserve((size_t) n*11*sizeof(double));
        partD_.reserve((size_t) std::max(partTotal_, 1)*10*sizeof(double));
        cgPartD_.reserve(kPRegions*kCgMaxBlocks*sizeof(double));
        cgD_.reserve(sizeof(CgState));
        doneD_.reserve((s_.max_iter+2)*sizeof(int));
        nbD_.reserve(n*sizeof(int));
        T4* pos = posT<T>();
        // the initial guess: from the history when there is one, then its first product rides on the pair setup
        const bool induce = polarizable_ && s_.polarizable;
        const GuessCoef gc = guessCoefficients(newStep);
        const bool dual = induce && gc.nc > 0;
        const size_t stride = 3*(size_t) nGlobal_;
        if (induce)
            histD_.reserve(kNHist*stride*sizeof(double));
        if (dual) {
            tb(kPhCg);
            launchGuess(st_, n, orderD_.as<int>(), gidxD_.as<int>(), histD_.as<double>(), stride, gc,
                        alphaSD_.as<double>(), bD_.as<double>(), xD_.as<double>());
            launchToVec4<T>(st_, n, xD_.as<double>(), pTD_.as<T4>(), rgT<T>());
            te();
        }
        // field of the charges and permanent dipoles (and T x0): compaction + pair setup (fused or as two
        // passes, the autotuner's choice) + PME
        PmeStreamScope pmeScope(*this, overlapPme());
        fork();
        pairSetup<T>(setupLanes(), fused_, dual ? pTD_.as<T4>() : nullptr);
        preReady_ = false;
        if (induce && preRadius_ > 0) {   // the preconditioner's short list, from the stored kernels
            srCountD_.reserve(n*sizeof(int));
            srJD_.reserve((size_t) n*kPreCap*sizeof(int));
            srGD_.reserve((size_t) n*kPreCap*sizeof(T2));
            srOverD_.reserve(sizeof(int));
            tb(kPhCg);
            launchZero<int>(st_, srOverD_.as<int>(), 1, nullptr);
            launchShortList<T>(st_, n, pos, candStartD_.as<int>(), nbD_.as<int>(), pjD_.as<int>(), g12D_.as<T2>(),
                               bt, (T) (preRadius_*preRadius_), srCountD_.as<int>(), srJD_.as<int>(),
                               srGD_.as<T2>(), srOverD_.as<int>());
            te();
            preReady_ = true;
        }
        pmeSolve<T>(permTD_.as<T4>(), true, pos, 0, false, erD_.as<double>(), nullptr, nullptr, nullptr);
        if (dual)   // the guess's reciprocal field (its direct field rides on the setup pass)
            pmeSolve<T>(pTD_.as<T4>(), false, pos, 0, false, trD_.as<double>(), nullptr, nullptr, nullptr);
        join();
        launchCombineB(st_, n, edD_.as<double>(), erD_.as<double>(), permD_.as<double>(), cself, bD_.as<double>());
        // induced dipoles
        int pushSlot = -1;
        if (induce) {
            if (!dual) {
                tb(kPhCg);
                launchGuess(st_, n, orderD_.as<int>(), gidxD_.as<int>(), histD_.as<double>(), stride, gc,
                            alphaSD_.as<double>(), bD_.as<double>(), xD_.as<double>());
                launchToVec4<T>(st_, n, xD_.as<double>(), pTD_.as<T4>(), rgT<T>());
                te();
                matvec<T>(pos, pTD_.as<T4>(), nullptr);
            }
            const int its = cgSolve<T>(pos, bD_.as<double>(), xD_.as<double>(), lastIterations_, residual_);
            if (its >= 0)   // (else enqueued ahead: known when the launch completes)
                iterations_ = its;
            const int h = (int) histSlots_.size();
            const bool push = !frozen_ && !noHistory_ && (newStep || h == 0);
            if (push)
                pushSlot = freeSlot();
            tb(kPhCg);
            launchPost<T>(st_, n, s_.peek, alphaSD_.as<double>(), rD_.as<double>(), xD_.as<double>(),
                          permD_.as<double>(), orderD_.as<int>(), gidxD_.as<int>(),
                          push ? histD_.as<double>()+pushSlot*stride : nullptr, dD_.as<double>(), rgSD_.as<double>(),
                          dTD_.as<T4>());
            te();
        }
        else {
            launchZero<double>(st_, xD_.as<double>(), 3*(size_t) n, nullptr);
            launchZero<double>(st_, rD_.as<double>(), 3*(size_t) n, nullptr);
            launchPost<T>(st_, n, 0.0, alphaSD_.as<double>(), rD_.as<double>(), xD_.as<double>(), permD_.as<double>(),
                          orderD_.as<int>(), gidxD_.as<int>(), nullptr, dD_.as<double>(), rgSD_.as<double>(),
                          dTD_.as<T4>());
            iterations_ = 0;
            residual_ = 0;
        }
        // final stage
        double* res = resD_.as<double>();
        fork();
        tb(kPhPairFinal);
        const GvdwArgs<T> ga = gvdwArgs<T>();
        launchPairFinal<T>(st_, n, pos, dTD_.as<T4>(), rgT<T>(), (T) s_.beta, candStartD_.as<int>(), nbD_.as<int>(),
                           pjD_.as<int>(), g12D_.as<T2>(), g03D_.as<T2>(), bt, fpairD_.as<double>(),
                           edD_.as<double>(), perAtomD_.as<double>(), wantStrain, ga, lanes_, nextListDirection());
        if (ga.gv && nSkip_ > 0)
            launchGvdwSkip<T>(st_, n, orderD_.as<int>(), rankD_.as<int>(), pos, rgT<T>(), skipStartD_.as<int>(),
                              skipPartD_.as<int>(), bt, (T) (s_.cutoff*s_.cutoff), wantStrain, ga,
                              fpairD_.as<double>(), perAtomD_.as<double>());
        systemSums(perAtomD_.as<double>(), wantStrain ? 10 : 1, res, kResPair);
        if (ga.gv)
            systemSums(gvdwAtomD_.as<double>(), 1, res, kResGvdw);
        te();
        pmeSolve<T>(dTD_.as<T4>(), true, pos, 1, wantStrain, erD_.as<double>(), frecD_.as<double>(), nullptr, res);
        join();
        tb(kPhFinal);
        launchFinalCombine(st_, n, dD_.as<double>(), xD_.as<double>(), alphaSD_.as<double>(), edD_.as<double>(),
                           erD_.as<double>(), cself, etotD_.as<double>(), perAtomD_.as<double>());
        systemSums(perAtomD_.as<double>(), 11, res, kResComb);
        launchCovalentForces(st_, n, rankD_.as<int>(), posDev_, covStartD_.as<int>(), covPartnerD_.as<int>(),
                             covStrengthD_.as<double>(), revStartD_.as<int>(), revEntryD_.as<int>(),
                             covAtomOfD_.as<int>(), etotD_.as<double>(), bd, fcovD_.as<double>(),
                             perAtomD_.as<double>());
        systemSums(perAtomD_.as<double>(), 9, res, kResCov);
        te();
        if (asyncEval_)
            pendingPush_ = pushSlot;   // committed when the launch completes valid
        else
            commitPush(pushSlot);
        haveResults_ = true;
    }

    void commitPush(int slot) {
        if (slot < 0)
            return;
        histSlots_.push_front(slot);
        while (histSlots_.size() > 4)
            histSlots_.pop_back();
    }

    // Milliseconds per CG product (field pass + PME solve of the total dipoles) with the PME beside the pair pass
    // or after it on the engine stream.
    template<class T>
    double timeProduct(bool overlap, int reps) {
        typedef typename Vec4<T>::type T4;
        hipEvent_t a = nextEvent(), b = nextEvent();
        const bool wasTiming = timing_;
        timing_ = false;
        adjointReady_ = false;   // the products overwrite the evaluation's CG buffers
        PmeStreamScope scope(*this, overlap);
        matvec<T>(posT<T>(), dTD_.as<T4>(), nullptr);   // warm-up
        HIPCHK(hipEventRecord(a, st_));
        for (int r = 0; r < reps; r++)
            matvec<T>(posT<T>(), dTD_.as<T4>(), nullptr);
        HIPCHK(hipEventRecord(b, st_));
        HIPCHK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        evUsed_ -= 2;
        timing_ = wasTiming;
        return ms/reps;
    }

    // Milliseconds per evaluation spent in the setup pass (with its compaction) at the given lane count.
    template<class T>
    double timeSetup(int lanes, bool fused, int reps) {
        typedef typename Vec4<T>::type T4;
        hipEvent_t a = nextEvent(), b = nextEvent();
        const bool wasTiming = timing_;
        timing_ = false;
        adjointReady_ = false;   // the timed passes overwrite the evaluation's pair-pass outputs
        pairSetup<T>(lanes, fused, pTD_.as<T4>());   // warm-up
        HIPCHK(hipEventRecord(a, st_));
        for (int r = 0; r < reps; r++)
            pairSetup<T>(lanes, fused, pTD_.as<T4>());
        HIPCHK(hipEventRecord(b, st_));
        HIPCHK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        evUsed_ -= 2;
        timing_ = wasTiming;
        return ms/reps;
    }

    // Milliseconds per evaluation spent in the pair passes at the given lane count (current lists and dipoles).
    template<class T>
    double timePairPasses(int lanes, bool fused, int reps, bool otf) {
        typedef typename Vec4<T>::type T4;
        typedef typename Vec2<T>::type T2;
        const int n = n_;
        const BoxSel<T> bt = boxSel<T>();
        T4* pos = posT<T>();
        hipEvent_t a = nextEvent(), b = nextEvent();
        const bool wasTiming = timing_;
        timing_ = false;
        adjointReady_ = false;   // the timed passes overwrite the evaluation's pair-pass outputs
        auto once = [&] {
            pairSetup<T>(lanes, fused, pTD_.as<T4>());
            for (int k = 0; k <= iterations_; k++)
                launchPairField<T>(st_, n, pos, dTD_.as<T4>(), candStartD_.as<int>(), nbD_.as<int>(), pjD_.as<int>(),
                                   g12D_.as<T2>(), bt, tdD_.as<double>(), nullptr, lanes, otf, (T) s_.beta,
                                   nextListDirection());
            launchPairFinal<T>(st_, n, pos, dTD_.as<T4>(), rgT<T>(), (T) s_.beta, candStartD_.as<int>(),
                               nbD_.as<int>(), pjD_.as<int>(), g12D_.as<T2>(), g03D_.as<T2>(), bt,
                               fpairD_.as<double>(), edD_.as<double>(), perAtomD_.as<double>(), false,
                               gvdwArgs<T>(), lanes, nextListDirection());
        };
        once();   // warm-up
        HIPCHK(hipEventRecord(a, st_));
        for (int r = 0; r < reps; r++)
            once();
        HIPCHK(hipEventRecord(b, st_));
        HIPCHK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        evUsed_ -= 2;
        timing_ = wasTiming;
        return ms/reps;
    }

    // Parameter gradients after an evaluation at the same positions (the sorted state, lists and stored kernels
    // of that evaluation are reused; the PME grids are recomputed).  gCaller: dL/dF (caller's order, device), or
    // null for the energy gradient alone; wE: dL/dE per system.  Outputs in the caller's order (device, each may
    // be null): gq, gR, galpha per atom, gc per covalent entry in input order.
    template<class T>
    void adjoint(const double* gCaller, const double* wE, double* gq, double* gR, double* galpha, double* gc) {
        typedef typename Vec4<T>::type T4;
        const int n = n_;
        const size_t v3 = 3*sizeof(double)*n;
        const BoxSel<double> bd = boxSel<double>();
        const BoxSel<T> bt = boxSel<T>();
        const double cself = 4.0*s_.beta*s_.beta*s_.beta/(3.0*kSqrtPi);
        const bool tan = gCaller != nullptr;
        const bool induce = polarizable_ && s_.polarizable;
        T4* pos = posT<T>();
        pairOutD_.reserve(10*sizeof(double)*n);
        pmeOutD_.reserve(10*sizeof(double)*n);
        for (DevBuf* b : {&gSD_, &pdotD_, &lamD_, &adjBD_, &ddotD_, &etotDotD_})
            b->reserve(v3);
        for (DevBuf* b : {&gTD_, &ddotTD_})
            b->reserve(n*sizeof(T4));
        if (tan)
            gridTD_.reserve((size_t) realTotal_*sizeof(T));
        if (batch_)
            upload(wED_, wE, sys_.size()*sizeof(double));
        // the potential / field of the charges and total dipoles (and their tangents for the displacement g and
        // the dipole change ddT): pair pass over the stored kernels, the PME grids and their tangent grids
        // (the PME part on the second stream beside the pair pass, as in the forward evaluation)
        PmeStreamScope pmeScope(*this, overlapPme());
        auto tangentPass = [&](const T4* ddT) {
            fork();
            launchPairTangent<T>(st_, n, pos, rgT<T>(), (T) s_.beta, dTD_.as<T4>(), ddT, ddT ? gTD_.as<T4>() : nullptr,
                                 candStartD_.as<int>(), nbD_.as<int>(), pjD_.as<int>(),
                                 g12D_.as<typename Vec2<T>::type>(), g03D_.as<typename Vec2<T>::type>(), bt,
                                 pairOutD_.as<double>());
            pmeGrids<T>(dTD_.as<T4>(), true, pos, 0, false, gridD_.as<T>(), nullptr, nullptr, -1);
            if (ddT)
                pmeGrids<T>(dTD_.as<T4>(), true, pos, 0, false, gridTD_.as<T>(), nullptr, nullptr, -1, ddT,
                            gTD_.as<T4>());
            launchGatherTangent<T>(pmeSt_, pmeSel(), n, pmeAtomD_.as<int>(), baseD_.as<int>(), splD_.as<T>(),
                                   gridD_.as<T>(), ddT ? gridTD_.as<T>() : nullptr, ddT ? gTD_.as<T4>() : nullptr,
                                   pmeOutD_.as<double>());
            join();
        };
        adjIterations_ = 0;
        adjResidual_ = 0;
        const double* lambda = nullptr;
        if (tan) {
            launchSort3(st_, n, orderD_.as<int>(), gCaller, gSD_.as<double>());
            launchToVec4<T>(st_, n, gSD_.as<double>(), gTD_.as<T4>());
            launchPermTangent(st_, n, orderD_.as<int>(), posDev_, gCaller, covStartD_.as<int>(),
                              covPartnerD_.as<int>(), covStrengthD_.as<double>(), bd, pdotD_.as<double>());
            launchDipoleTangent<T>(st_, n, pdotD_.as<double>(), nullptr, ddotD_.as<double>(), ddotTD_.as<T4>());
            if (induce) {
                // lambda = (alpha^-1 - T)^-1 (-E'(g)) at fixed induced dipoles
                tangentPass(ddotTD_.as<T4>());
                launchAdjointRhs(st_, n, alphaSD_.as<double>(), pairOutD_.as<double>(), pmeOutD_.as<double>(),
                                 pdotD_.as<double>(), cself, adjBD_.as<double>());
                launchGuess(st_, n, orderD_.as<int>(), gidxD_.as<int>(), nullptr, 0, GuessCoef{},
                            alphaSD_.as<double>(), adjBD_.as<double>(), lamD_.as<double>());
                launchToVec4<T>(st_, n, lamD_.as<double>(), pTD_.as<T4>(), rgT<T>());
                matvec<T>(pos, pTD_.as<T4>(), nullptr);
                adjIterations_ = cgSolve<T>(pos, adjBD_.as<double>(), lamD_.as<double>(), lastAdjIterations_,
                                            adjResidual_);
                lambda = lamD_.as<double>();
                launchDipoleTangent<T>(st_, n, pdotD_.as<double>(), lambda, ddotD_.as<double>(), ddotTD_.as<T4>());
            }
            tangentPass(ddotTD_.as<T4>());
        }
        else
            tangentPass(nullptr);
        ParamAtomArgs a{};
        a.n = n;
        a.order = orderD_.as<int>();
        a.q = qSD_.as<double>();
        a.alpha = alphaSD_.as<double>();
        a.mu = xD_.as<double>();
        a.lambda = lambda;
        a.d = dD_.as<double>();
        a.ddot = tan ? ddotD_.as<double>() : nullptr;
        a.pairOut = pairOutD_.as<double>();
        a.pmeOut = pmeOutD_.as<double>();
        a.ke = s_.ke;
        a.wE = wE[0];
        a.cself = cself;
        a.qSelf = 2.0*s_.beta/kSqrtPi;
        a.qBackground = kPi*sys_[0].Qtot/(sys_[0].volume*s_.beta*s_.beta);
        a.qBackSys = batch_ ? qBackD_.as<double>() : nullptr;
        a.wESys = batch_ ? wED_.as<double>() : nullptr;
        a.sysOf = batch_ ? sysOfD_.as<int>() : nullptr;
        a.gq = gq;
        a.galpha = galpha;
        a.gR = gR;
        a.etot = etotD_.as<double>();
        a.etotDot = etotDotD_.as<double>();
        launchParamAtoms(st_, a);
        if (gc)
            launchParamCov(st_, covN_, covAtomOfD_.as<int>(), covPartnerD_.as<int>(), covInputD_.as<int>(),
                           rankD_.as<int>(), posDev_, gCaller, etotD_.as<double>(), etotDotD_.as<double>(), bd, s_.ke,
                           wE[0], batch_ ? wED_.as<double>() : nullptr, gc);
    }

    void assemble(bool wantStrain) {
        const double* rAll = resH_.as<double>();
        const double ke = s_.ke;
        const double cself = 4.0*s_.beta*s_.beta*s_.beta/(3.0*kSqrtPi);
        for (size_t s = 0; s < sys_.size(); s++) {
            Sys& y = sys_[s];
            const double* r = rAll+s*kResCount;
            y.terms.direct = ke*r[kResPair];
            y.terms.reciprocal = ke*r[kResRec];
            y.terms.self = ke*(-(s_.beta/kSqrtPi)*y.q2-0.5*cself*r[kResComb]);
            const double bg = -kPi*y.Qtot*y.Qtot/(2.0*y.volume*s_.beta*s_.beta);
            y.terms.background = ke*bg;
            y.terms.polarization = ke*r[kResComb+1];
            y.terms.gvdw = gvdwOn_ ? r[kResGvdw] : 0.0;
            if (wantStrain) {
                for (int k = 0; k < 9; k++)
                    y.strain[k] = r[kResPair+1+k]+r[kResRec+1+k]+r[kResComb+2+k]+r[kResCov+k];
                for (int a = 0; a < 3; a++)
                    y.strain[4*a] -= bg;
                for (int k = 0; k < 9; k++)
                    y.strain[k] *= ke;
            }
        }
    }
};

}   // namespace

}   // namespace pgmhip

// ------------------------------------------------------------------------------------------------ C ABI

using pgmhip::DeviceEngine;
using pgmhip::PosSource;

struct pgm_engine {
    std::unique_ptr<DeviceEngine> eng;
};

namespace {
thread_local std::string g_createError;
}   // namespace

namespace pgmhip {
// The message pgm_last_error(NULL) returns: failures of calls without an engine.
void setGlobalError(const char* msg) { g_createError = msg; }
}   // namespace pgmhip

namespace {

template<class Fn>
// settle: first complete a launch still in flight (every call but launch, finish and wait)
int guarded(pgm_engine* e, Fn fn, bool settle = true) {
    if (!e || !e->eng)
        return -2;
    try {
        if (settle)
            e->eng->settle();
        fn(*e->eng);
        e->eng->lastError.clear();
        return 0;
    }
    catch (const std::exception& ex) {
        e->eng->lastError = ex.what();
        return -1;
    }
}

void requireSingle(const DeviceEngine& g) {
    if (g.numSystems() > 1)
        throw std::invalid_argument("pgm_hip: the engine holds a batch of systems: use the batch entry points");
}
}   // namespace

extern "C" {

void pgm_default_settings(pgm_settings* s) {
    if (!s)
        return;
    s->cutoff = 0.9;
    s->beta = 3.5;
    s->grid[0] = s->grid[1] = s->grid[2] = 36;
    s->order = 6;
    s->tol = 1e-5;
    s->max_iter = 100;
    s->peek = 0.65;
    s->predictor = 2;
    s->ke = 138.935456;
    s->device = 0;
    s->skin = 0.15;
    s->precision = PGM_MIXED;
    s->polarizable = 1;
}

pgm_engine* pgm_create(const pgm_settings* s) {
    if (!s) {
        g_createError = "pgm_hip: pgm_create needs settings";
        return nullptr;
    }
    try {
        pgm_engine* e = new pgm_engine;
        e->eng.reset(new DeviceEngine(*s));
        return e;
    }
    catch (const std::exception& ex) {
        g_createError = ex.what();
        return nullptr;
    }
}

void pgm_destroy(pgm_engine* e) { delete e; }

const char* pgm_last_error(const pgm_engine* e) {
    if (!e || !e->eng)
        return g_createError.c_str();
    return e->eng->lastError.c_str();
}

int pgm_set_atoms(pgm_engine* e, int n, const double* q, const double* radius, const double* alpha,
                  const int* global_index, int n_global) {
    return guarded(e, [&](DeviceEngine& g) { g.setAtoms(n, q, radius, alpha, global_index, n_global); });
}

int pgm_set_gvdw(pgm_engine* e, int repulsion, double rvdw, const double* sqrt_a, const double* sqrt_c6,
                 const double* b, int nskip, const int* skip_i, const int* skip_j) {
    return guarded(e, [&](DeviceEngine& g) { g.setGvdw(repulsion, rvdw, sqrt_a, sqrt_c6, b, nskip, skip_i, skip_j); });
}

int pgm_set_candidates(pgm_engine* e, int nseg, const int* seg_i, const int* seg_shift, const int* seg_begin,
                       const int* cand_j) {
    return guarded(e, [&](DeviceEngine& g) { g.setCandidates(nseg, seg_i, seg_shift, seg_begin, cand_j); });
}

int pgm_set_preconditioner(pgm_engine* e, double radius, int jacobi_self) {
    return guarded(e, [&](DeviceEngine& g) { g.setPreconditioner(radius, jacobi_self != 0); });
}

int pgm_set_covalent(pgm_engine* e, int ncov, const int* atom, const int* partner, const double* strength) {
    return guarded(e, [&](DeviceEngine& g) { g.setCovalent(ncov, atom, partner, strength); });
}

int pgm_compute(pgm_engine* e, const double* pos, const double* box, int new_step, double* f, double* strain,
                pgm_terms* terms) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.evaluate(PosSource::HostF64, pos, box, new_step != 0, f, nullptr, nullptr, 0, 0.0, strain, terms, nullptr,
                   false);
    });
}

int pgm_compute_f32(pgm_engine* e, const float* pos, const double* box, int new_step, double* strain,
                    pgm_terms* terms) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.evaluate(PosSource::HostF32, pos, box, new_step != 0, nullptr, nullptr, nullptr, 0, 0.0, strain, terms,
                   nullptr, true);
    });
}

int pgm_forces(pgm_engine* e, double* out) {
    return guarded(e, [&](DeviceEngine& g) {
        const std::vector<double>& f = g.forces();
        if (f.empty())
            throw std::logic_error("pgm_hip: no forces: pgm_compute_f32 has not run");
        std::copy(f.begin(), f.end(), out);
    });
}

int pgm_compute_device(pgm_engine* e, const void* posq4, const double* box, int new_step, long long* force_fixed,
                       int stride, double fixed_scale, void* stream, pgm_terms* terms) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.evaluate(PosSource::DeviceF4, posq4, box, new_step != 0, nullptr, nullptr, force_fixed, stride, fixed_scale,
                   nullptr, terms, static_cast<hipStream_t>(stream), false);
    });
}

int pgm_launch_device(pgm_engine* e, const void* posq4, const double* box, int new_step, long long* force_fixed,
                      int stride, double fixed_scale, void* stream) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.launchDevice(posq4, box, new_step != 0, force_fixed, stride, fixed_scale, static_cast<hipStream_t>(stream));
    }, false);
}

int pgm_finish_device(pgm_engine* e, pgm_terms* terms, int* valid) {
    return guarded(e, [&](DeviceEngine& g) {
        const bool ok = g.finishDevice();
        if (terms)
            *terms = g.lastTerms();
        if (valid)
            *valid = ok ? 1 : 0;
    }, false);
}

int pgm_wait_device(pgm_engine* e, void* stream) {
    return guarded(e, [&](DeviceEngine& g) { g.waitDevice(static_cast<hipStream_t>(stream)); }, false);
}

int pgm_launch_pending(const pgm_engine* e) { return (e && e->eng && e->eng->launchPending()) ? 1 : 0; }

int pgm_compute_device_f64(pgm_engine* e, const double* pos_dev, const double* box, int new_step, double* f_dev,
                           double* strain, void* stream, pgm_terms* terms) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.evaluate(PosSource::DeviceF64, pos_dev, box, new_step != 0, nullptr, f_dev, nullptr, 0, 0.0, strain, terms,
                   static_cast<hipStream_t>(stream), false);
    });
}

int pgm_compute_device_f64_noalloc(pgm_engine* e, const double* pos_dev, const double* box, double* f_dev,
                                   double* strain, void* stream, pgm_terms* terms) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        if (!pos_dev || !box)
            throw std::invalid_argument("pgm_hip: compute needs positions and a box");
        g.withDevice([&] { g.evaluateNoAlloc(pos_dev, box, f_dev, strain, terms, static_cast<hipStream_t>(stream)); });
    });
}

int pgm_field(pgm_engine* e, const double* pos, const double* box, const double* q, const double* d, double* field) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.field(pos, box, q, d, field);
    });
}

int pgm_iterations(const pgm_engine* e) { return (e && e->eng) ? e->eng->iterations() : -1; }
double pgm_residual(const pgm_engine* e) { return (e && e->eng) ? e->eng->residual() : -1.0; }

int pgm_induced_dipoles(pgm_engine* e, double* mu) {
    return guarded(e, [&](DeviceEngine& g) { g.dipoles(true, mu); });
}

int pgm_permanent_dipoles(pgm_engine* e, double* p) {
    return guarded(e, [&](DeviceEngine& g) { g.dipoles(false, p); });
}

int pgm_reset_history(pgm_engine* e) {
    return guarded(e, [&](DeviceEngine& g) { g.resetHistory(); });
}

long pgm_list_builds(const pgm_engine* e) { return (e && e->eng) ? e->eng->listBuilds() : -1; }
long long pgm_num_candidates(const pgm_engine* e) { return (e && e->eng) ? e->eng->numCandidates() : -1; }
long long pgm_num_pairs(pgm_engine* e) {
    long long v = -1;
    if (guarded(e, [&](DeviceEngine& g) { v = g.numPairs(); }) != 0)
        return -1;
    return v;
}

long pgm_late_list_checks(const pgm_engine* e) { return (e && e->eng) ? e->eng->lateListChecks() : -1; }

int pgm_list_check_state(const pgm_engine* e, double* out) {
    if (!e || !e->eng || !out)
        return -2;
    e->eng->listCheckState(out);
    return 0;
}

int pgm_set_history(pgm_engine* e, int nset, const double* mu_sets) {
    return guarded(e, [&](DeviceEngine& g) { g.setHistory(nset, mu_sets); });
}

int pgm_backward(pgm_engine* e, const double* pos, const double* box, const double* dl_dforces, double dl_denergy,
                 double* grad_q, double* grad_radius, double* grad_alpha, double* grad_cov) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.backward(PosSource::HostF64, pos, box, dl_dforces, false, &dl_denergy, grad_q, grad_radius, grad_alpha,
                   grad_cov, false, nullptr);
    });
}

int pgm_backward_device(pgm_engine* e, const double* pos_dev, const double* box, const double* dl_dforces_dev,
                        double dl_denergy, double* grad_q_dev, double* grad_radius_dev, double* grad_alpha_dev,
                        double* grad_cov_dev, void* stream) {
    return guarded(e, [&](DeviceEngine& g) {
        requireSingle(g);
        g.backward(PosSource::DeviceF64, pos_dev, box, dl_dforces_dev, true, &dl_denergy, grad_q_dev, grad_radius_dev,
                   grad_alpha_dev, grad_cov_dev, true, static_cast<hipStream_t>(stream));
    });
}

int pgm_adjoint_iterations(const pgm_engine* e) { return (e && e->eng) ? e->eng->adjointIterations() : -1; }
double pgm_adjoint_residual(const pgm_engine* e) { return (e && e->eng) ? e->eng->adjointResidual() : -1.0; }

int pgm_set_batch(pgm_engine* e, int nsys, const int* atom_offsets, const int* grids) {
    return guarded(e, [&](DeviceEngine& g) { g.setSystems(nsys, atom_offsets, grids); });
}

int pgm_num_systems(const pgm_engine* e) { return (e && e->eng) ? e->eng->numSystems() : -1; }

int pgm_compute_batch(pgm_engine* e, const double* pos, const double* boxes, int new_step, double* f,
                      pgm_terms* terms) {
    return guarded(e, [&](DeviceEngine& g) {
        g.evaluate(PosSource::HostF64, pos, boxes, new_step != 0, f, nullptr, nullptr, 0, 0.0, nullptr, terms,
                   nullptr, false);
    });
}

int pgm_compute_batch_device(pgm_engine* e, const double* pos_dev, const double* boxes, int new_step, double* f_dev,
                             void* stream, pgm_terms* terms) {
    return guarded(e, [&](DeviceEngine& g) {
        g.evaluate(PosSource::DeviceF64, pos_dev, boxes, new_step != 0, nullptr, f_dev, nullptr, 0, 0.0, nullptr,
                   terms, static_cast<hipStream_t>(stream), false);
    });
}

int pgm_backward_batch(pgm_engine* e, const double* pos, const double* boxes, const double* dl_dforces,
                       const double* dl_denergy, double* grad_q, double* grad_radius, double* grad_alpha,
                       double* grad_cov) {
    return guarded(e, [&](DeviceEngine& g) {
        std::vector<double> wE(g.numSystems(), 0.0);
        if (dl_denergy)
            std::copy(dl_denergy, dl_denergy+g.numSystems(), wE.begin());
        g.backward(PosSource::HostF64, pos, boxes, dl_dforces, false, wE.data(), grad_q, grad_radius, grad_alpha,
                   grad_cov, false, nullptr);
    });
}

int pgm_backward_batch_device(pgm_engine* e, const double* pos_dev, const double* boxes,
                              const double* dl_dforces_dev, const double* dl_denergy, double* grad_q_dev,
                              double* grad_radius_dev, double* grad_alpha_dev, double* grad_cov_dev, void* stream) {
    return guarded(e, [&](DeviceEngine& g) {
        std::vector<double> wE(g.numSystems(), 0.0);
        if (dl_denergy)
            std::copy(dl_denergy, dl_denergy+g.numSystems(), wE.begin());
        g.backward(PosSource::DeviceF64, pos_dev, boxes, dl_dforces_dev, true, wE.data(), grad_q_dev,
                   grad_radius_dev, grad_alpha_dev, grad_cov_dev, true, static_cast<hipStream_t>(stream));
    });
}

int pgm_autotune(pgm_engine* e, int reps, int* lanes, int* fused, double* ms) {
    return guarded(e, [&](DeviceEngine& g) {
        const int best = g.autotune(reps);
        if (lanes)
            *lanes = best;
        if (fused)
            *fused = g.fused() ? 1 : 0;
        if (ms)
            std::copy(g.tuneTimes(), g.tuneTimes()+6, ms);
    });
}

int pgm_set_lanes(pgm_engine* e, int lanes, int fused) {
    return guarded(e, [&](DeviceEngine& g) { g.setLanes(lanes, fused != 0); });
}

int pgm_setup_lanes(const pgm_engine* e) { return (e && e->eng) ? e->eng->setupLanes() : -1; }
int pgm_set_setup_lanes(pgm_engine* e, int lanes) {
    return guarded(e, [&](DeviceEngine& g) { g.setSetupLanes(lanes); });
}

int pgm_set_field_otf(pgm_engine* e, int on) {
    return guarded(e, [&](DeviceEngine& g) { g.setFieldOtf(on != 0); });
}
int pgm_field_otf(const pgm_engine* e) { return (e && e->eng) ? (e->eng->fieldOtf() ? 1 : 0) : -1; }

int pgm_set_pme_overlap(pgm_engine* e, int on) {
    return guarded(e, [&](DeviceEngine& g) { g.setPmeOverlap(on != 0); });
}
int pgm_pme_overlap(const pgm_engine* e) { return (e && e->eng) ? (e->eng->pmeOverlap() ? 1 : 0) : -1; }
int pgm_pme_overlap_times(pgm_engine* e, double* ms) {
    return guarded(e, [&](DeviceEngine& g) {
        if (!ms)
            throw std::invalid_argument("pgm_hip: pgm_pme_overlap_times needs ms[2]");
        std::copy(g.overlapTimes(), g.overlapTimes()+2, ms);
    });
}

int pgm_set_timing(pgm_engine* e, int on) {
    return guarded(e, [&](DeviceEngine& g) { g.setTiming(on != 0); });
}

int pgm_set_timing_phases(pgm_engine* e, unsigned mask) {
    return guarded(e, [&](DeviceEngine& g) { g.setTiming(mask != 0, mask); });
}

int pgm_get_timing(pgm_engine* e, double* ms, long* counts, int nphases) {
    return guarded(e, [&](DeviceEngine& g) {
        if (nphases < pgmhip::kNumPhases)
            throw std::invalid_argument("pgm_hip: pgm_get_timing needs room for every phase");
        g.timing(ms, counts);
    });
}

int pgm_freeze_history(pgm_engine* e, int frozen) {
    return guarded(e, [&](DeviceEngine& g) { g.freezeHistory(frozen != 0); });
}

}   // extern "C"
