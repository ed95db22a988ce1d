// The qcg_context object shared by the host translation units.
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>

#include "../../include/qcg.h"
#include "eri_launch.h"
#include "onee.h"
#include "onee_launch.h"
#include "xc_launch.h"

using namespace qcg;

struct qcg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool has_basis = false;
    HostBasis basis;
    int precision = QCG_PREC_FP64;
    double fp32_cutoff = 1e-6;
    double schwarz_thr = kSchwarzDefault;
    int charge = 0;
    // device tables
    double* d_rys = nullptr;
    float* d_rys_f = nullptr;
    // pairs
    std::vector<PairRec> pairs;  // grouped by class, sorted by Q desc within a class
    int class_off[kNumPairClasses + 1] = {0};
    PairRec* d_pairs = nullptr;
    PrimPair<double>* d_pp = nullptr;
    PrimPair<float>* d_ppf = nullptr;
    std::vector<double> Qshell;  // nshell x nshell (single molecule)
    // batch of molecules: nmol = 1 for qcg_set_basis
    int nmol = 1, max_nbf = 0;
    std::vector<HostBasis> mols;             // per-molecule basis (local atom / bf indices)
    std::vector<int> mol_nbf_h;
    std::vector<int64_t> mat_off_h;          // offset of each molecule's nbf^2 block
    std::vector<int> ket_off_h, ket_cnt_h;   // [class][mol], relative to the class start
    // contraction buckets of each (class, molecule) list: pairs are ordered by surviving
    // primitive-pair count (descending), then Schwarz bound (descending); bucket_bnd[class][mol]
    // holds the bucket boundaries 0 = b_0 < b_1 < ... < b_n = count (positions in the list)
    std::vector<std::vector<int>> bucket_bnd;
    int* d_mol_nbf = nullptr;
    int64_t* d_mat_off = nullptr;
    int* d_ket_off = nullptr;
    int* d_ket_cnt = nullptr;
    int64_t mat_total = 0;
    // work buffers
    double *d_D = nullptr, *d_J = nullptr, *d_K = nullptr;
    // J and K accumulators before symmetrisation, fixed-point (fixacc.h): one word per element
    fix_t *d_Jacc = nullptr, *d_Kacc = nullptr;
    // their scales per molecule (4 doubles each), recomputed at the start of every build from the
    // shell-pair Schwarz matrices (qsh, packed per molecule), the shell of every function and D
    double *d_fix = nullptr, *d_qsh = nullptr, *d_rowsum = nullptr, *d_qmax = nullptr;
    int64_t* d_qsh_off = nullptr;
    int *d_mol_ns = nullptr, *d_bf_shell = nullptr, *d_bf_off = nullptr, *d_row_mol = nullptr;
    int *d_shell_first = nullptr, *d_shell_first_off = nullptr;
    FixData fixdata{};
    double *d_H = nullptr, *d_F = nullptr;
    std::vector<void*> owned;
    std::vector<PrimPair<double>> h_pp;  // host copy of the primitive pairs (flop model counts)
    // graph replay of the device Fock build (launch sequence captured once per key)
    bool use_graph = false;
    hipGraph_t fock_graph = nullptr;
    hipGraphExec_t fock_exec = nullptr;
    double fock_key[10] = {};
    // ev_done (below) is recorded after every launch of the exec, so waiting on it before the
    // exec is destroyed covers a launch still in flight (without touching the caller's stream,
    // which may no longer exist)
    void drop_fock_graph() {
        if (fock_exec && ev_done) (void)hipEventSynchronize(ev_done);
        if (fock_exec) (void)hipGraphExecDestroy(fock_exec);
        if (fock_graph) (void)hipGraphDestroy(fock_graph);
        fock_exec = nullptr;
        fock_graph = nullptr;
    }
    // Stream order of the entry points that share the accumulators (d_Jacc, d_Kacc, d_grad) and
    // the fork/join events: each one, on its stream st, first waits for the previous one's work
    // (ev_done, recorded on last_stream) and records ev_done at its end, so calls issued on
    // different streams of one context run one after the other on the device.
    hipEvent_t ev_done = nullptr;
    hipStream_t last_stream = nullptr;
    // per-iteration trace of the last qcg_rhf_scf / qcg_ks_scf (qcg_scf_history)
    std::vector<qcg_scf_iteration> scf_history;
    // per-class profiling of the J/K build: events around each class launch
    bool profiling = false;
    hipEvent_t ev[2 * 21 * 2 + 2] = {};
    bool ev_made = false;
    bool ev_recorded[21][2] = {};
    // J/K work plan: items per (class pair, precision part), rebuilt when options or tiles change
    struct Tile { int block = 64, reps = 1; };
    Tile tile[21][2];  // per quartet class and precision part (0 FP64, 1 FP32)
    int class_nq[21] = {};  // Cartesian components per quartet class
    // mixed precision, per class: run its FP64 and FP32 items in one launch (jk_*_mp kernels, with
    // the FP64 part's tile) or in two; qcg_autotune_refine may split a class
    bool fuse_cls[21] = {true, true, true, true, true, true, true, true, true, true, true,
                         true, true, true, true, true, true, true, true, true, true};
    // bra pairs with more primitive pairs than this are cut into primitive ranges (cut_items_split)
    int prim_split = 0;
    // launch schedule of the class kernels: classes launch longest first onto the least-loaded
    // auxiliary stream, by the cost per class set by qcg_autotune or qcg_set_class_cost
    // (class_cost_set), else by plan_cost, the work plan's model flops per class
    double class_ms[21] = {};
    bool class_cost_set = false;
    double plan_cost[21] = {};
    int launch_order[21] = {}, launch_stream[21] = {};
    // ERI gradient work items (built once per threshold and primitive split) and the device
    // gradient (natom x 3)
    int4* d_grad_items = nullptr;
    int2* d_grad_segs = nullptr;
    int grad_off[21] = {}, grad_cnt[21] = {};
    double grad_key[2] = {-1.0, -1.0};
    double* d_grad = nullptr;
    double* d_W = nullptr;  // energy-weighted density of the gradient (nbf x nbf)
    // device basis copy and work list of the one-electron gradient kernel
    int* d_onee_int = nullptr;
    double* d_onee_dbl = nullptr;
    int4* d_onee_work = nullptr;
    int onee_nwork = 0;
    BasisView onee_view{};
    std::vector<int4> items;  // J/K work items over the ket segments segs (cut_items)
    std::vector<int2> segs;
    int item_off[21][2] = {}, item_cnt[21][2] = {};
    int bra_pp[kNumPairClasses] = {};  // per pair class: most primitive pairs of one item's bra
    int4* d_items = nullptr;
    int2* d_segs = nullptr;
    double plan_key[4] = {-1.0, -1.0, -1.0, -1.0};
    int plan_sig = -1, tile_sig = 0;
#ifndef QCG_STREAMS
#define QCG_STREAMS 4
#endif
    static constexpr int kStreams = QCG_STREAMS;
    hipStream_t aux[kStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kStreams] = {};
    // the KS Fock build's XC stream, concurrent with the J/K class kernels (made on first use)
    hipStream_t xc_stream = nullptr;
    hipEvent_t ev_xc_fork = nullptr, ev_xc_join = nullptr;
    // cached class counts / flop model, valid for (precision, fp32_cutoff, schwarz_thr) of stats_key
    std::vector<qcg_class_stat> stats;
    // exchange-correlation grid and per-point buffers
    rocblas_handle blas = nullptr;
    int64_t G = 0;
    std::vector<double> grid_xyz, grid_w;
    std::vector<int> grid_atom;
    double *d_pts = nullptr, *d_w = nullptr, *d_phi = nullptr, *d_dphx = nullptr, *d_dphy = nullptr,
           *d_dphz = nullptr, *d_X = nullptr, *d_aow = nullptr, *d_V = nullptr, *d_exc = nullptr,
           *d_Vparts = nullptr;  // split-K partial products of V = Phi^T aow
    int v_chunks = 0, v_chunks_mfma = 0, v_chunks_mfma32 = 0;
    // XC build (qcg_set_xc_mode): 0 = FP32 MFMA kernels for the mixed option, FP64 MFMA kernels up
    // to 352 functions and rocBLAS DGEMMs around point_kernel above; 1 = rocBLAS DGEMMs for both;
    // 2 = MFMA kernels for both
    int xc_mode = 0;
    // FP32 copies of Phi and grad Phi plus FP32 work arrays for the mixed-precision XC build,
    // made on its first use for the current grid
    float *d_phi32 = nullptr, *d_aow32 = nullptr, *d_D32 = nullptr;
    float4* d_grid4 = nullptr;  // (Phi, dPhi/dx, dPhi/dy, dPhi/dz) per element, FP32
    GridShell* d_gshell = nullptr;
    double *d_gexp = nullptr, *d_gcn = nullptr;
    void free_grid() {
        for (void** p : {(void**)&d_pts, (void**)&d_w, (void**)&d_phi, (void**)&d_dphx, (void**)&d_dphy,
                         (void**)&d_dphz, (void**)&d_X, (void**)&d_aow, (void**)&d_V, (void**)&d_exc, (void**)&d_Vparts,
                         (void**)&d_gshell, (void**)&d_gexp, (void**)&d_gcn, (void**)&d_phi32, (void**)&d_grid4,
                         (void**)&d_aow32, (void**)&d_D32}) {
            release(*p);
            *p = nullptr;
        }
        G = 0;
    }
    double stats_key[3] = {-1.0, -1.0, -1.0};

    template <typename T>
    hipError_t alloc(T** p, size_t n) {
        hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) owned.push_back((void*)*p);
        return e;
    }
    void release(void* p) {
        if (!p) return;
        auto it = std::find(owned.begin(), owned.end(), p);
        if (it != owned.end()) owned.erase(it);
        (void)hipFree(p);
    }
    void free_basis() {
        drop_fock_graph();
        for (void** p : {(void**)&d_pairs, (void**)&d_pp, (void**)&d_ppf, (void**)&d_D, (void**)&d_J, (void**)&d_K,
                         (void**)&d_Jacc, (void**)&d_Kacc, (void**)&d_fix, (void**)&d_qsh, (void**)&d_rowsum,
                         (void**)&d_qsh_off, (void**)&d_mol_ns, (void**)&d_bf_shell, (void**)&d_bf_off,
                         (void**)&d_qmax, (void**)&d_row_mol, (void**)&d_shell_first, (void**)&d_shell_first_off,
                         (void**)&d_H, (void**)&d_F, (void**)&d_mol_nbf,
                         (void**)&d_mat_off, (void**)&d_ket_off, (void**)&d_ket_cnt, (void**)&d_items,
                         (void**)&d_segs, (void**)&d_grad_items, (void**)&d_grad_segs, (void**)&d_grad,
                         (void**)&d_W, (void**)&d_onee_int, (void**)&d_onee_dbl, (void**)&d_onee_work}) {
            release(*p);
            *p = nullptr;
        }
        has_basis = false;
        plan_sig = -1;
        free_grid();
    }
};


namespace qcg {

inline int fail(qcg_context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

// True when st is inside an open stream capture (then nothing may allocate, copy synchronously or
// create runtime objects: such calls would end the capture with an error).
inline bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Order the work about to be enqueued on st after the context's previous device work, then
// (stream_done) mark its end.  No-ops inside a capture: a captured sequence is ordered when its
// graph is launched.
inline hipError_t stream_begin(qcg_context* c, hipStream_t st) {
    if (capturing(st)) return hipSuccess;
    if (!c->ev_done) {
        hipError_t e = hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    if (c->last_stream && c->last_stream != st) return hipStreamWaitEvent(st, c->ev_done, 0);
    return hipSuccess;
}
inline hipError_t stream_done(qcg_context* c, hipStream_t st) {
    if (capturing(st) || !c->ev_done) return hipSuccess;
    c->last_stream = st;
    return hipEventRecord(c->ev_done, st);
}

// Graph, exec and timing events of one captured launch sequence, released on every exit path
// (after the stream the exec was launched on has drained).
struct GraphHold {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    GraphHold() = default;
    GraphHold(const GraphHold&) = delete;
    GraphHold& operator=(const GraphHold&) = delete;
    ~GraphHold() {
        if (st) (void)hipStreamSynchronize(st);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Integral-direct J/K into Jacc/Kacc then J = Jacc + Jacc^T, K = Kacc + Kacc^T (dJ/dK may be null).
// acc (optional) replaces the context's accumulators, scales and row-sum scratch (the no-allocation
// entry point passes its static scratch); no_alloc: fail instead of (re)building the plan;
// symmetrize = false: leave J and K in the accumulators (a Fock build reads them there).
struct JkAcc {
    fix_t* Jacc;
    fix_t* Kacc;
    double* fix;
    double* rowsum;
};
int run_jk(qcg_context* c, const double* dD, double* dJ, double* dK, hipStream_t st, const JkAcc* acc = nullptr,
           bool no_alloc = false, bool symmetrize = true);
// The work plan, auxiliary streams and events exist for the current options and tiles.
bool jk_prepared(const qcg_context* c);
// Looks up this device's address of the no-allocation entry point's static scratch (noalloc.hip);
// once per device, with the device current.
hipError_t noalloc_setup(int device);
// Everything run_jk needs that allocates or creates runtime objects (work-item plan, auxiliary
// streams, fork/join events), done ahead so that run_jk itself can be captured into a graph.
int prepare_jk(qcg_context* c);
// ERI gradient at the device density dD, ADDED into dgrad on st; returns after the launches.
// dW != nullptr: also the one-electron terms sum D dH/dR - W dS/dR (onee_grad_kernel).
int eri_grad_launch(qcg_context* c, const double* dD, double hf_frac, double* dgrad, hipStream_t st,
                    const double* dW = nullptr);
// Device basis and work list of the one-electron gradient kernel (once per basis).
int onee_setup(qcg_context* c);
// S, T, V of the basis B (any output may be null), nbf x nbf row-major host arrays.
int one_electron_device(qcg_context* c, const HostBasis& B, double* S, double* T, double* V);
// The same for n bases in one go (S, T, V: arrays of n host pointers, or null).
int one_electron_device_many(qcg_context* c, int n, const HostBasis* const* B, double* const* S, double* const* T,
                             double* const* V);
// W = 2 sum_i eps_i C_mu,i C_nu,i over the n_occ occupied orbitals (C row-major nbf x nbf)
std::vector<double> energy_weighted_density(int nbf, int n_occ, const double* C, const double* eps);
// eri_grad_launch for the host density D (and energy-weighted density W, optional) into the
// context's zeroed d_grad, on c->stream
int eri_grad_start(qcg_context* c, const double* D, double hf_frac, const double* W = nullptr);
// waits on c->stream and copies d_grad to out
int grad_finish(qcg_context* c, double* out);
double radial_norm(double a, int l);

// First position in [lo, hi) of the list starting at pairs[base] with Qp * Q < thr (Q descending
// inside [lo, hi)).
inline int prefix_end(const qcg_context* c, int base, int lo, int hi, double Qp, double thr) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Qp * c->pairs[base + mid].Q >= thr) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The surviving kets of bra pair p (position in molecule m's class-cb list) in molecule m's
// class-ck list, one call f(begin, split, end) per contraction bucket: [begin, end) are the kets
// with Q_P Q_Q >= thr (the reference's screen, integrals/eri/mod.rs:1008-1023), [begin, split)
// those with Q_P Q_Q >= cut (cut <= thr: split == end).  For cb == ck only kets at positions <= p
// (each unordered pair of pairs once).  Positions are relative to the molecule's class-ck list.
template <typename F>
void for_ket_ranges(const qcg_context* c, int cb, int ck, int m, int p, double thr, double cut, F&& f) {
    const int nmol = c->nmol;
    const int base = c->class_off[ck] + c->ket_off_h[ck * nmol + m];
    const double Qp = c->pairs[c->class_off[cb] + c->ket_off_h[cb * nmol + m] + p].Q;
    const std::vector<int>& bnd = c->bucket_bnd[ck * nmol + m];
    for (size_t b = 0; b + 1 < bnd.size(); ++b) {
        const int bs = bnd[b];
        int be = bnd[b + 1];
        if (cb == ck) {
            if (bs > p) break;
            be = std::min(be, p + 1);
        }
        const int e = prefix_end(c, base, bs, be, Qp, thr);
        if (e == bs) continue;
        const int sp = cut > thr ? prefix_end(c, base, bs, e, Qp, cut) : e;
        f(bs, sp, e);
    }
}

#define HIPCHK(ctx, expr)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(ctx, QCG_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)


}  // namespace qcg
