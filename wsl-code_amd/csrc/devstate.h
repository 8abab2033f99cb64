// Device-side state of a solver (shared by the HIP translation units):
// operator store, Krylov vectors, per-instance recurrences and the
// multigrid hierarchy.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "precond.h"
#include "solver.h"

namespace vemd {

struct MultiState;
struct MgBuild;

template <typename T>
struct DevOps {
    T* O = nullptr;
    T* L = nullptr;
    T* scale = nullptr;
    T* minv = nullptr;
    GroupDesc* groups = nullptr;
};

template <typename T>
struct DevKrylov {
    T* R[2] = {nullptr, nullptr};
    T* Z[2] = {nullptr, nullptr};
    T* WB[2] = {nullptr, nullptr};
    T* x = nullptr;
    T* Az = nullptr;  // cell-owned and multiplier rows of the apply
    T* H = nullptr;   // edge results per cell side (summed in the Krylov sweep)
    T* b = nullptr;
    hipGraphExec_t graph = nullptr;   // check_every iterations
    hipGraphExec_t graph_tail = nullptr;  // a few iterations, launched near convergence
    int graph_iters = 0;
    bool graph_amg = false;
};

struct DeviceState {
    int device = 0;
    hipStream_t stream = nullptr;
    // apply variant chosen per valence (index ne, -1: no cells of that valence)
    int tuned64[16], tuned32[16];
    float tuned_ms64[16], tuned_ms32[16];  // the chosen variant's time on the valence's largest group
    DevOps<double> o64;
    DevOps<float> o32;
    uint32_t* idx = nullptr;
    uint32_t* wmask = nullptr;
    int* hside = nullptr;  // cell-side slots of every edge (two per edge)
    int* tile_group = nullptr;
    int* tile_inst = nullptr;  // batch instance of every tile
    int* cell_of_gpos = nullptr;
    int ntiles = 0;
    double* pdot = nullptr;
    double* plam = nullptr;
    double* p1 = nullptr;       // per-chunk partials of the Krylov sweeps
    double* p2 = nullptr;
    int nblk = 0;               // grid of the plain elementwise kernels
    int ninst = 1;
    VecChunk* chunks = nullptr;
    int nchunks = 0;
    int* pchunks = nullptr;
    int npchunks = 0;
    int* inst_tile0 = nullptr;
    int* inst_chunk0 = nullptr;
    MState* st = nullptr;       // per instance
    MState* st_host = nullptr;  // pinned copy
    int* all_done = nullptr;
    double* rtol = nullptr;     // per-instance inner tolerances
    double* rtol_host = nullptr;  // pinned
    double* dsum = nullptr;     // per-instance sums
    double* hsum = nullptr;     // pinned
    DevKrylov<double> k64;
    DevKrylov<float> k32;
    double* x64 = nullptr;  // outer solution
    double* r64 = nullptr;  // outer residual
    double* y64 = nullptr;
    double* b64 = nullptr;
    double* bprob = nullptr;  // the problem's own right-hand side, internal order
    float* r32 = nullptr;
    double* h64 = nullptr;  // cell-side scratch of the plain applies
    float* h32 = nullptr;
    AmgDev<double>* amg64 = nullptr;  // multigrid pressure preconditioner (when built)
    AmgDev<float>* amg32 = nullptr;
    MgBuild* mgb = nullptr;  // index data of the device rebuild of the multigrid values (made on first use)
    double* q1 = nullptr;  // per-chunk partials of the pressure-block dots (zero outside it)
    double* q2 = nullptr;
    std::vector<void*> allocs;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // Order of the calls that use the solver's shared device scratch (side
    // slots, partial sums, Krylov and multi-right-hand-side buffers): recorded
    // at the end of every apply or solve on whatever stream it ran; the next
    // call's stream waits for it.  ev_caller marks a caller's stream at the
    // start of a device-resident solve.
    hipEvent_t ev_order = nullptr, ev_caller = nullptr;
    MultiState* multi = nullptr;  // multi-right-hand-side solver state (allocated on first use)
};

// One apply or solve: its stream (work) waits for the previous user of the
// scratch and, given a caller stream, for the work queued there; on leaving,
// even by an exception, it records itself as the last user, and the caller's
// stream waits for it.  Calls that overlap on one solver therefore run one
// after the other on the device, whichever streams they were given.
struct ScratchUse {
    DeviceState& d;
    hipStream_t work, caller;
    bool with_caller;
    ScratchUse(DeviceState& dd, hipStream_t w, bool wc = false, hipStream_t c = nullptr)
        : d(dd), work(w), caller(c), with_caller(wc) {
        if (with_caller) {
            HIPCHK(hipEventRecord(d.ev_caller, caller));
            HIPCHK(hipStreamWaitEvent(work, d.ev_caller, 0));
        }
        HIPCHK(hipStreamWaitEvent(work, d.ev_order, 0));
    }
    ~ScratchUse() {
        (void)hipEventRecord(d.ev_order, work);
        if (with_caller) (void)hipStreamWaitEvent(caller, d.ev_order, 0);
    }
    ScratchUse(const ScratchUse&) = delete;
    ScratchUse& operator=(const ScratchUse&) = delete;
};

inline void* dalloc(DeviceState& d, size_t bytes) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
    d.allocs.push_back(p);
    return p;
}

template <typename T>
inline T* dupload(DeviceState& d, const std::vector<T>& v) {
    T* p = (T*)dalloc(d, v.size() * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}


// Multi-right-hand-side path (MFMA element apply, simultaneous MINRES).
void multi_free(DeviceState& d);
void multi_invalidate(DeviceState& d);  // drops the captured iteration graph

}  // namespace vemd
