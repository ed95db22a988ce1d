// The qcg_context object shared by the host translation units.
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>

#include "../../include/qcg.h"
#include "eri_launch.h"
#include "onee.h"
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
    int* d_mol_nbf = nullptr;
    int64_t* d_mat_off = nullptr;
    int* d_ket_off = nullptr;
    int* d_ket_cnt = nullptr;
    int64_t mat_total = 0;
    // work buffers
    double *d_D = nullptr, *d_J = nullptr, *d_K = nullptr, *d_Jacc = nullptr, *d_Kacc = nullptr;
    double *d_H = nullptr, *d_F = nullptr;
    std::vector<void*> owned;
    std::vector<PrimPair<double>> h_pp;  // host copy of the primitive pairs (flop model counts)
    // per-class profiling of the J/K build: events around each class launch
    bool profiling = false;
    hipEvent_t ev[2 * 21 * 2 + 2] = {};
    bool ev_made = false;
    bool ev_recorded[21][2] = {};
    // J/K work plan: items per (class pair, precision part), rebuilt when options or tiles change
    struct Tile { int block = 64, reps = 1; };
    Tile tile[21];
    int class_nq[21] = {};  // Cartesian components per quartet class
    std::vector<int4> items;
    int item_off[21][2] = {}, item_cnt[21][2] = {};
    int4* d_items = nullptr;
    double plan_key[4] = {-1.0, -1.0, -1.0, -1.0};
    int plan_sig = -1, tile_sig = 0;
    static constexpr int kStreams = 4;
    hipStream_t aux[kStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kStreams] = {};
    // cached class counts / flop model, valid for (precision, fp32_cutoff, schwarz_thr) of stats_key
    std::vector<qcg_class_stat> stats;
    // exchange-correlation grid and per-point buffers
    rocblas_handle blas = nullptr;
    int64_t G = 0;
    std::vector<double> grid_xyz, grid_w;
    std::vector<int> grid_atom;
    double *d_pts = nullptr, *d_w = nullptr, *d_phi = nullptr, *d_dphx = nullptr, *d_dphy = nullptr,
           *d_dphz = nullptr, *d_X = nullptr, *d_aow = nullptr, *d_V = nullptr, *d_exc = nullptr;
    GridShell* d_gshell = nullptr;
    double *d_gexp = nullptr, *d_gcn = nullptr;
    void free_grid() {
        for (void** p : {(void**)&d_pts, (void**)&d_w, (void**)&d_phi, (void**)&d_dphx, (void**)&d_dphy,
                         (void**)&d_dphz, (void**)&d_X, (void**)&d_aow, (void**)&d_V, (void**)&d_exc,
                         (void**)&d_gshell, (void**)&d_gexp, (void**)&d_gcn}) {
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
        for (void** p : {(void**)&d_pairs, (void**)&d_pp, (void**)&d_ppf, (void**)&d_D, (void**)&d_J, (void**)&d_K,
                         (void**)&d_Jacc, (void**)&d_Kacc, (void**)&d_H, (void**)&d_F, (void**)&d_mol_nbf,
                         (void**)&d_mat_off, (void**)&d_ket_off, (void**)&d_ket_cnt, (void**)&d_items}) {
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

// Integral-direct J/K into Jacc/Kacc then J = Jacc + Jacc^T, K = Kacc + Kacc^T (dJ/dK may be null).
int run_jk(qcg_context* c, const double* dD, double* dJ, double* dK, hipStream_t st);
// Everything run_jk needs that allocates or creates runtime objects (work-item plan, auxiliary
// streams, fork/join events), done ahead so that run_jk itself can be captured into a graph.
int prepare_jk(qcg_context* c);
double radial_norm(double a, int l);

#define HIPCHK(ctx, expr)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(ctx, QCG_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)


}  // namespace qcg
