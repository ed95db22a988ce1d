// Host side of the exchange-correlation path: grid set-up, V_xc / E_xc builds, KS-SCF.
//
// Per build (build_vxc_gga_matmul, dft/ks_scf.rs:1150-1296, with compute_chi_d :936 and
// compute_density_and_gradient_matmul :1028):
//   X = Phi D                       (GEMM, G x nbf x nbf)
//   point_kernel: rho, grad rho, sigma, B3LYP, aow row, E_xc
//   V = Phi^T aow, V_xc = (V + V^T)/2   (GEMM, nbf x nbf x G)
// Phi and grad Phi are evaluated once per grid and kept resident in HBM.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"
#include "grid.h"
#include "scf.h"

using namespace qcg;

namespace {

int ensure_blas(qcg_context* c) {
    if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success)
        return fail(c, QCG_DEVICE, "rocblas_create_handle failed");
    return QCG_OK;
}

constexpr int64_t kVChunk = 16384;
// Largest basis whose default FP64 XC build runs on the MFMA kernels: up to here the density kernel
// holds all of X's columns per wavefront (16 x 22 MFMA column tiles); (H2O)16 (304 functions):
// 11.5 ms per FP64 XC build against 12.1 ms for the library DGEMMs around point_kernel
constexpr int kXcMfmaFp64MaxNbf = 352;

int upload_grid(qcg_context* c) {
    const HostBasis& B = c->basis;
    c->free_grid();
    const int64_t G = (int64_t)c->grid_w.size();
    c->G = G;
    if (G == 0) return QCG_OK;
    const size_t gn = (size_t)G * B.nbf;
    HIPCHK(c, c->alloc(&c->d_pts, 3 * (size_t)G));
    HIPCHK(c, c->alloc(&c->d_w, (size_t)G));
    HIPCHK(c, c->alloc(&c->d_phi, gn));
    HIPCHK(c, c->alloc(&c->d_dphx, gn));
    HIPCHK(c, c->alloc(&c->d_dphy, gn));
    HIPCHK(c, c->alloc(&c->d_dphz, gn));
    HIPCHK(c, c->alloc(&c->d_X, gn));
    HIPCHK(c, c->alloc(&c->d_aow, gn));
    HIPCHK(c, c->alloc(&c->d_V, (size_t)B.nbf * B.nbf));
    // V = Phi^T aow has a long inner dimension (the grid): split it into chunks of about
    // kVChunk points, one strided-batched GEMM over the chunks plus a sum, so the GEMM has
    // chunks x tiles workgroups instead of a handful of tiles with a 10^6-long K loop
    c->v_chunks = (int)std::max<int64_t>(1, G / kVChunk);
    // the MFMA V kernel: grid chunks for two full rounds of (chunk, output tile) blocks at the
    // kernel's occupancy, at least 1024 points per chunk, and in FP32 (whose accumulators sum a
    // chunk) at most 4096 points per chunk.  (H2O)16: FP64 160-wide tiles at one block per CU,
    // 128 chunks of 8 200 points (2048 / 1024 / 512 blocks: 11.70 / 11.57 / 11.50 ms per FP64
    // XC build); FP32 at two blocks per CU, 256 chunks of 4 100 points (512 blocks were 1 % faster
    // with twice the FP32 summation length).
    {
        int cus = 0;
        HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        const int64_t most = std::max<int64_t>(1, G / 1024);
        auto chunks = [&](bool fp32) {
            const int64_t blocks = 2LL * std::max(cus, 1) * xc_v_blocks_per_cu(B.nbf, fp32);
            int64_t n = std::max<int64_t>(1, blocks / xc_mfma_vtiles(B.nbf, fp32));
            if (fp32) n = std::max<int64_t>(n, (G + 4095) / 4096);
            return (int)std::min(n, most);
        };
        c->v_chunks_mfma = chunks(false);
        c->v_chunks_mfma32 = chunks(true);
    }
    HIPCHK(c, c->alloc(&c->d_Vparts, (size_t)std::max({c->v_chunks, c->v_chunks_mfma, c->v_chunks_mfma32}) * B.nbf *
                                         B.nbf));
    HIPCHK(c, c->alloc(&c->d_exc, 1));
    HIPCHK(c, hipMemcpy(c->d_pts, c->grid_xyz.data(), 3 * (size_t)G * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_w, c->grid_w.data(), (size_t)G * sizeof(double), hipMemcpyHostToDevice));
    // shells for the basis evaluation: coef * (2a/pi)^(3/4) (4a)^(l/2) per primitive
    std::vector<GridShell> gs(B.nshell);
    std::vector<double> ex, cn;
    for (int s = 0; s < B.nshell; ++s) {
        GridShell& g = gs[s];
        g.cx = B.centers[3 * s];
        g.cy = B.centers[3 * s + 1];
        g.cz = B.centers[3 * s + 2];
        g.l = B.l[s];
        g.nprim = B.nprim[s];
        g.poff = (int)ex.size();
        g.bf = B.bf_off[s];
        g.atom = B.shell_atom[s];
        g.pad = 0;
        double amin = 1e300;
        for (int p = 0; p < B.nprim[s]; ++p) {
            const double a = B.exps[B.prim_off[s] + p];
            amin = std::min(amin, a);
            ex.push_back(a);
            cn.push_back(B.coefs[B.prim_off[s] + p] * radial_norm(a, B.l[s]));
        }
        g.amin = amin;
    }
    HIPCHK(c, c->alloc(&c->d_gshell, gs.size()));
    HIPCHK(c, c->alloc(&c->d_gexp, ex.size()));
    HIPCHK(c, c->alloc(&c->d_gcn, cn.size()));
    HIPCHK(c, hipMemcpy(c->d_gshell, gs.data(), gs.size() * sizeof(GridShell), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_gexp, ex.data(), ex.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_gcn, cn.data(), cn.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, launch_basis_on_grid((int)G, B.nshell, c->d_gshell, c->d_gexp, c->d_gcn, c->d_pts, B.nbf, c->d_phi,
                                   c->d_dphx, c->d_dphy, c->d_dphz, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QCG_OK;
}

// FP32 copies of Phi and grad Phi (and FP32 work arrays) for the mixed-precision XC build
int ensure_xc32(qcg_context* c, hipStream_t st) {
    if (c->d_phi32 && c->d_grid4 && c->d_aow32 && c->d_D32) return QCG_OK;
    const size_t gn = (size_t)c->G * c->basis.nbf, nn = (size_t)c->basis.nbf * c->basis.nbf;
    HIPCHK(c, c->alloc(&c->d_phi32, gn));
    HIPCHK(c, c->alloc(&c->d_grid4, gn));
    HIPCHK(c, c->alloc(&c->d_aow32, gn));
    HIPCHK(c, c->alloc(&c->d_D32, nn));
    HIPCHK(c, launch_to_f32(c->d_phi, c->d_phi32, gn, st));
    HIPCHK(c, launch_to_f32x4(c->d_phi, c->d_dphx, c->d_dphy, c->d_dphz, c->d_grid4, gn, st));
    return QCG_OK;
}

// V_xc and E_xc with the MFMA kernels (xc_mfma.hip): FP32 products in mixed precision, FP64 otherwise
int xc_build_mfma(qcg_context* c, const double* dD, double* dV, double* dExc, hipStream_t st) {
    const int nbf = c->basis.nbf, G = (int)c->G;
    const bool mixed = c->precision == QCG_PREC_MIXED;
    const int nch = xc_mfma_chunks(G, mixed ? c->v_chunks_mfma32 : c->v_chunks_mfma);
    HIPCHK(c, launch_fill(dExc, 1, 0.0, st));
    if (mixed) {
        int rc = ensure_xc32(c, st);
        if (rc) return rc;
        HIPCHK(c, launch_to_f32(dD, c->d_D32, (size_t)nbf * nbf, st));
        HIPCHK(c, launch_xc_mfma(G, nbf, c->v_chunks_mfma32, c->d_w, c->d_phi32, c->d_grid4, c->d_D32, c->d_aow32, dExc,
                                 c->d_Vparts, st));
    } else {
        HIPCHK(c, launch_xc_mfma(G, nbf, c->v_chunks_mfma, c->d_w, c->d_phi, c->d_dphx, c->d_dphy, c->d_dphz, dD, c->d_aow,
                                 dExc, c->d_Vparts, st));
    }
    HIPCHK(c, launch_sum_parts(c->d_Vparts, nch, (size_t)nbf * nbf, c->d_V, st));
    HIPCHK(c, launch_sym_half(c->d_V, dV, nbf, st));
    if (mixed)
        HIPCHK(c, launch_exc_d32_correction(dV, dD, c->d_D32, (size_t)nbf * nbf, dExc, st));
    return QCG_OK;
}

int xc_build(qcg_context* c, const double* dD, double* dV, double* dExc, hipStream_t st) {
    if (c->G <= 0) return fail(c, QCG_NO_GRID, "no grid: call qcg_build_grid or qcg_set_grid");
    // mode 0: the MFMA kernels (FP32 products for the mixed option, FP64 otherwise), except FP64
    // above kXcMfmaFp64MaxNbf functions, where the density kernel makes column passes and the
    // library DGEMMs win; mode 1: library DGEMMs for both; mode 2: MFMA kernels for both
    const bool mfma = c->xc_mode == 2 ||
                      (c->xc_mode == 0 && (c->precision == QCG_PREC_MIXED || c->basis.nbf <= kXcMfmaFp64MaxNbf));
    if (mfma) return xc_build_mfma(c, dD, dV, dExc, st);
    int rc = ensure_blas(c);
    if (rc) return rc;
    const int nbf = c->basis.nbf;
    const rocblas_int G = (rocblas_int)c->G;
    if (rocblas_set_stream(c->blas, st) != rocblas_status_success) return fail(c, QCG_DEVICE, "rocblas_set_stream");
    const double one = 1.0, zero = 0.0;
    // X^T (nbf x G, column-major) = D (nbf x nbf) * Phi^T (nbf x G)
    if (rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nbf, G, nbf, &one, dD, nbf, c->d_phi,
                      nbf, &zero, c->d_X, nbf) != rocblas_status_success)
        return fail(c, QCG_DEVICE, "rocblas_dgemm (Phi D)");
    HIPCHK(c, launch_fill(dExc, 1, 0.0, st));
    HIPCHK(c, launch_xc_points((int)G, nbf, c->d_w, c->d_phi, c->d_dphx, c->d_dphy, c->d_dphz, c->d_X, c->d_aow, dExc,
                               nullptr, st));
    // V (column-major nbf x nbf) = Phi^T (nbf x G) * aow^T^T, split over grid chunks: chunk k
    // covers points [k Gc, (k+1) Gc), the last one also the remainder
    const int nch = c->v_chunks;
    const rocblas_int Gc = G / nch, Glast = G - (nch - 1) * Gc;
    const rocblas_stride sg = (rocblas_stride)Gc * nbf, sv = (rocblas_stride)nbf * nbf;
    if (nch > 1 &&
        rocblas_dgemm_strided_batched(c->blas, rocblas_operation_none, rocblas_operation_transpose, nbf, nbf, Gc, &one,
                                      c->d_phi, nbf, sg, c->d_aow, nbf, sg, &zero, c->d_Vparts, nbf, sv,
                                      nch - 1) != rocblas_status_success)
        return fail(c, QCG_DEVICE, "rocblas_dgemm_strided_batched (Phi^T aow)");
    if (rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_transpose, nbf, nbf, Glast, &one,
                      c->d_phi + (size_t)(nch - 1) * sg, nbf, c->d_aow + (size_t)(nch - 1) * sg, nbf, &zero,
                      c->d_Vparts + (size_t)(nch - 1) * sv, nbf) != rocblas_status_success)
        return fail(c, QCG_DEVICE, "rocblas_dgemm (Phi^T aow, last chunk)");
    HIPCHK(c, launch_sum_parts(c->d_Vparts, nch, (size_t)nbf * nbf, c->d_V, st));
    HIPCHK(c, launch_sym_half(c->d_V, dV, nbf, st));
    return QCG_OK;
}

// XC nuclear gradient at the device density dD, ADDED into dgrad (natom x 3) on stream st.
int xc_gradient(qcg_context* c, const double* dD, double* dgrad, hipStream_t st) {
    if (c->G <= 0) return fail(c, QCG_NO_GRID, "no grid: call qcg_build_grid or qcg_set_grid");
    int rc = ensure_blas(c);
    if (rc) return rc;
    const HostBasis& B = c->basis;
    const int nbf = B.nbf;
    const rocblas_int G = (rocblas_int)c->G;
    if (rocblas_set_stream(c->blas, st) != rocblas_status_success) return fail(c, QCG_DEVICE, "rocblas_set_stream");
    const double one = 1.0, zero = 0.0;
    double *dM = nullptr, *dt = nullptr;
    HIPCHK(c, c->alloc(&dM, (size_t)G * nbf));
    HIPCHK(c, c->alloc(&dt, 3 * (size_t)G));
    rc = QCG_OK;
    if (rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nbf, G, nbf, &one, dD, nbf, c->d_phi,
                      nbf, &zero, c->d_X, nbf) != rocblas_status_success)
        rc = fail(c, QCG_DEVICE, "rocblas_dgemm (Phi D)");
    hipError_t e = hipSuccess;
    if (!rc) e = launch_xc_grad_points((int)G, nbf, c->d_w, c->d_phi, c->d_dphx, c->d_dphy, c->d_dphz, c->d_X,
                                       c->d_aow, dt, st);
    if (!rc && e != hipSuccess) rc = fail(c, QCG_DEVICE, std::string("xc_grad_points: ") + hipGetErrorString(e));
    // M (= A D, row-major G x nbf) as column-major nbf x G: D^T (= D) * A^T
    if (!rc && rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nbf, G, nbf, &one, dD, nbf,
                             c->d_aow, nbf, &zero, dM, nbf) != rocblas_status_success)
        rc = fail(c, QCG_DEVICE, "rocblas_dgemm (A D)");
    if (!rc) e = launch_xc_grad_accumulate((int)G, B.nshell, c->d_gshell, c->d_gexp, c->d_gcn, c->d_pts, nbf, c->d_X,
                                           dM, dt, dgrad, st);
    if (!rc && e != hipSuccess) rc = fail(c, QCG_DEVICE, std::string("xc_grad_accumulate: ") + hipGetErrorString(e));
    const hipError_t es = hipStreamSynchronize(st);  // the temporaries are read until here
    c->release(dM);
    c->release(dt);
    if (rc) return rc;
    HIPCHK(c, es);
    return QCG_OK;
}

}  // namespace

extern "C" {

int qcg_xc_gradient(qcg_context* c, const double* D, double* grad) {
    if (!c || !D || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "gradient: one molecule per context");
    HIPCHK(c, hipSetDevice(c->device));
    const int nbf = c->basis.nbf, n3 = 3 * c->basis.natom;
    double* dg = nullptr;
    HIPCHK(c, c->alloc(&dg, n3));
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, (size_t)nbf * nbf * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_fill(dg, n3, 0.0, c->stream));
    int rc = xc_gradient(c, c->d_D, dg, c->stream);
    if (!rc) {
        HIPCHK(c, hipMemcpyAsync(grad, dg, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->release(dg);
    return rc;
}

int qcg_ks_gradient(qcg_context* c, const double* D, const double* C, const double* eps, int n_occ, double* grad) {
    if (!c || !D || !C || !eps || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->G <= 0) return fail(c, QCG_NO_GRID, "no grid: call qcg_build_grid or qcg_set_grid");
    const int nbf = c->basis.nbf, n3 = 3 * c->basis.natom;
    if (n_occ < 0 || n_occ > nbf) return fail(c, QCG_INVALID_ARG, "n_occ");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "gradient: one molecule per context");
    HIPCHK(c, hipSetDevice(c->device));
    const std::vector<double> W = energy_weighted_density(nbf, n_occ, C, eps);
    // ERI part (B3lyp::A = 0.20 exact exchange) and one-electron part on the device, then the XC
    // part into the same device gradient; nuclear repulsion on the host
    int rc = eri_grad_start(c, D, 0.20, W.data());
    if (rc) return rc;
    rc = xc_gradient(c, c->d_D, c->d_grad, c->stream);
    std::vector<double> g2(n3, 0.0);
    nuclear_repulsion_gradient(c->basis, g2.data());
    const int rf = grad_finish(c, grad);
    if (rc) return rc;
    if (rf) return rf;
    for (int i = 0; i < n3; ++i) grad[i] += g2[i];
    return QCG_OK;
}

namespace {
// build_becke_grid's partition step on the device of the context
int device_partition(void* user, int64_t np, const double* pts, const int* owner, int nat, const double* axyz,
                     const double* dist, double* part) {
    qcg_context* c = static_cast<qcg_context*>(user);
    double *dp = nullptr, *da = nullptr, *dd = nullptr, *dw = nullptr;
    int* dow = nullptr;
    int rc = QCG_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == QCG_OK) rc = fail(c, QCG_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(c->alloc(&dp, 3 * (size_t)np), "hipMalloc");
    chk(c->alloc(&dow, (size_t)np), "hipMalloc");
    chk(c->alloc(&da, 3 * (size_t)nat), "hipMalloc");
    chk(c->alloc(&dd, (size_t)nat * nat), "hipMalloc");
    chk(c->alloc(&dw, (size_t)np), "hipMalloc");
    if (rc == QCG_OK) {
        chk(hipMemcpyAsync(dp, pts, 3 * (size_t)np * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload");
        chk(hipMemcpyAsync(dow, owner, (size_t)np * sizeof(int), hipMemcpyHostToDevice, c->stream), "upload");
        chk(hipMemcpyAsync(da, axyz, 3 * (size_t)nat * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload");
        chk(hipMemcpyAsync(dd, dist, (size_t)nat * nat * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload");
        chk(launch_becke_partition(np, dp, dow, nat, da, dd, dw, c->stream), "becke_partition");
        chk(hipMemcpyAsync(part, dw, (size_t)np * sizeof(double), hipMemcpyDeviceToHost, c->stream), "download");
        chk(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    }
    for (void* q : {(void*)dp, (void*)dow, (void*)da, (void*)dd, (void*)dw}) c->release(q);
    return rc;
}
}  // namespace

int qcg_build_grid(qcg_context* c, int n_radial, int quality, int pruning, int64_t* npts) {
    if (!c) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (n_radial <= 0 || (quality != 0 && quality != 1)) return fail(c, QCG_INVALID_ARG, "grid config");
    HIPCHK(c, hipSetDevice(c->device));
    GridConfig cfg;
    cfg.n_radial = n_radial;
    cfg.quality = quality;
    cfg.pruning = pruning ? 1 : 0;
    int rc = build_becke_grid(c->basis, cfg, c->grid_xyz, c->grid_w, c->grid_atom, device_partition, c);
    if (rc) return rc == QCG_DEVICE ? rc : fail(c, rc, "grid construction failed");
    rc = upload_grid(c);
    if (rc) return rc;
    if (npts) *npts = c->G;
    return QCG_OK;
}

int qcg_set_grid(qcg_context* c, int64_t npts, const double* xyz, const double* w) {
    if (!c || npts < 0 || (npts > 0 && (!xyz || !w))) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    c->grid_xyz.assign(xyz, xyz + 3 * npts);
    c->grid_w.assign(w, w + npts);
    c->grid_atom.assign(npts, -1);
    return upload_grid(c);
}

int qcg_get_grid(qcg_context* c, double* xyz, double* w, int* atom_idx) {
    if (!c) return QCG_INVALID_ARG;
    if (xyz) std::memcpy(xyz, c->grid_xyz.data(), c->grid_xyz.size() * sizeof(double));
    if (w) std::memcpy(w, c->grid_w.data(), c->grid_w.size() * sizeof(double));
    if (atom_idx) std::memcpy(atom_idx, c->grid_atom.data(), c->grid_atom.size() * sizeof(int));
    return QCG_OK;
}

int qcg_xc_eval(qcg_context* c, int64_t n, const double* rho, const double* sigma, double* exc, double* vrho,
                double* vsigma) {
    if (!c || n < 0 || (n > 0 && (!rho || !sigma || !exc || !vrho || !vsigma))) return QCG_INVALID_ARG;
    if (n == 0) return QCG_OK;
    HIPCHK(c, hipSetDevice(c->device));
    double* d = nullptr;
    HIPCHK(c, c->alloc(&d, 5 * (size_t)n));
    HIPCHK(c, hipMemcpyAsync(d, rho, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d + n, sigma, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_xc_eval((int)n, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, c->stream));
    HIPCHK(c, hipMemcpyAsync(exc, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(vrho, d + 3 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(vsigma, d + 4 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->release(d);
    return QCG_OK;
}

int qcg_set_xc_mode(qcg_context* c, int mode) {
    if (!c) return QCG_INVALID_ARG;
    if (mode < 0 || mode > 2) return fail(c, QCG_INVALID_ARG, "xc mode must be 0, 1 or 2");
    c->xc_mode = mode;
    return QCG_OK;
}

int qcg_xc_device(qcg_context* c, const double* dD, double* dV, double* dExc, void* stream) {
    if (!c || !dD || !dV || !dExc) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    return xc_build(c, dD, dV, dExc, stream ? (hipStream_t)stream : c->stream);
}

int qcg_xc(qcg_context* c, const double* D, double* Vxc, double* Exc) {
    if (!c || !D || !Vxc || !Exc) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n2 = (size_t)c->basis.nbf * c->basis.nbf;
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = xc_build(c, c->d_D, c->d_F, c->d_exc, c->stream);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(Vxc, c->d_F, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Exc, c->d_exc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QCG_OK;
}

}  // extern "C"

namespace {

struct KsUser {
    qcg_context* c;
    const double* H;
    double a;  // exact-exchange fraction (B3lyp::A = 0.20)
    double e_xc, e_j, e_k, e_1e;
};

// F = H + J - (a/2) K + V_xc;  E_elec = Tr DH + 1/2 Tr DJ - 1/4 a Tr DK + E_xc  (dft/ks_scf.rs:1960-1981)
int ks_fock_cb(void* u, const double* D, double* F, double* e_elec) {
    KsUser* k = (KsUser*)u;
    qcg_context* c = k->c;
    const int nbf = c->basis.nbf;
    const size_t n2 = (size_t)nbf * nbf;
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // V_xc on a stream of its own, concurrent with the J/K class kernels (the two share no buffer
    // and both only read D): the vector-ALU-bound integrals and the MFMA / HBM-bound grid products
    // overlap ((H2O)16: 9.80 -> 9.36 ms per mixed build, 16.35 -> 15.58 ms FP64)
    if (!c->xc_stream) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->xc_stream, hipStreamNonBlocking));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_xc_fork, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_xc_join, hipEventDisableTiming));
    }
    HIPCHK(c, hipEventRecord(c->ev_xc_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->xc_stream, c->ev_xc_fork, 0));
    int rc = run_jk(c, c->d_D, c->d_J, c->d_K, c->stream);
    if (!rc) rc = xc_build(c, c->d_D, c->d_F, c->d_exc, c->xc_stream);
    // join on every path: the next build's D upload must wait for this XC build
    HIPCHK(c, hipEventRecord(c->ev_xc_join, c->xc_stream));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_xc_join, 0));
    if (rc) return rc;
    std::vector<double> J(n2), K(n2), V(n2);
    double exc = 0.0;
    HIPCHK(c, hipMemcpyAsync(J.data(), c->d_J, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(K.data(), c->d_K, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(V.data(), c->d_F, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&exc, c->d_exc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tdh = 0.0, tdj = 0.0, tdk = 0.0;
    for (size_t i = 0; i < n2; ++i) {
        F[i] = k->H[i] + J[i] - 0.5 * k->a * K[i] + V[i];
        tdh += D[i] * k->H[i];
        tdj += D[i] * J[i];
        tdk += D[i] * K[i];
    }
    k->e_1e = tdh;
    k->e_j = 0.5 * tdj;
    k->e_k = -0.5 * k->a * 0.5 * tdk;
    k->e_xc = exc;
    *e_elec = tdh + 0.5 * tdj - 0.25 * k->a * tdk + exc;
    return QCG_OK;
}

// mixed precision: the last SCF iterations run FP64 builds (scf_driver's refine step)
int ks_refine_fp64(void* u) {
    static_cast<KsUser*>(u)->c->precision = QCG_PREC_FP64;
    return QCG_OK;
}
// the two-electron part's bound (exact-exchange fraction a)
int ks_bound_cb(void* u, const double* D, double* b) {
    KsUser* k = static_cast<KsUser*>(u);
    return qcg_error_bound(k->c, D, k->a, b);
}

}  // namespace

extern "C" int qcg_ks_scf(qcg_context* c, const qcg_scf_config* cfg, const double* D_init, qcg_scf_result* res,
                          double* D, double* C, double* eps, double* F) {
    if (!c || !cfg) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->G <= 0) return fail(c, QCG_NO_GRID, "no grid: call qcg_build_grid or qcg_set_grid");
    const HostBasis& B = c->basis;
    int nelec = -c->charge;
    for (int z : B.Z) nelec += z;
    if (nelec <= 0 || nelec % 2) return fail(c, QCG_INVALID_ARG, "closed-shell KS needs an even electron count");
    const size_t n2 = (size_t)B.nbf * B.nbf;
    std::vector<double> S(n2), T(n2), V(n2), H(n2);
    HIPCHK(c, hipSetDevice(c->device));
    const int r1 = one_electron_device(c, B, S.data(), T.data(), V.data());
    if (r1) return r1;
    for (size_t i = 0; i < n2; ++i) H[i] = T[i] + V[i];
    KsUser u{c, H.data(), 0.20, 0, 0, 0, 0};
    qcg_scf_result r{};
    const int prec = c->precision;
    int rc = scf_driver(B.nbf, nelec / 2, S.data(), H.data(), nuclear_repulsion(B), *cfg, ks_fock_cb, &u, D_init, &r,
                        D, C, eps, F, prec == QCG_PREC_MIXED ? ks_refine_fp64 : nullptr, ks_bound_cb, &c->scf_history);
    c->precision = prec;
    r.energy_xc = u.e_xc;
    r.energy_j = u.e_j;
    r.energy_k = u.e_k;
    r.energy_1e = u.e_1e;
    if (res) *res = r;
    if (rc == QCG_NOT_CONVERGED) return fail(c, rc, "ScfError::NotConverged");
    if (rc == QCG_NUMERICAL) return fail(c, rc, "ScfError::SingularOverlap");
    return rc;
}
