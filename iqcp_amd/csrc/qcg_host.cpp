// Host side of the qcg library: context, basis/pair setup, C ABI (include/qcg.h).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "context.h"
#include "rys.h"
#include "scf.h"

double qcg::radial_norm(double a, int l) {
    const double t = 2.0 * a / kPi;
    const double base = std::sqrt(std::sqrt(t * t * t));
    const double fa = 4.0 * a;
    return base * (l == 0 ? 1.0 : l == 1 ? std::sqrt(fa) : l == 2 ? fa : std::pow(fa, l / 2.0));
}

double qcg::nuclear_repulsion(const HostBasis& B) {
    // compute_nuclear_repulsion (basis/mod.rs:375-391)
    double e = 0.0;
    for (int i = 0; i < B.natom; ++i)
        for (int j = i + 1; j < B.natom; ++j) {
            const double dx = B.xyz[3 * i] - B.xyz[3 * j], dy = B.xyz[3 * i + 1] - B.xyz[3 * j + 1],
                         dz = B.xyz[3 * i + 2] - B.xyz[3 * j + 2];
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 1e-10) e += (double)B.Z[i] * B.Z[j] / r;
        }
    return e;
}

extern "C" {

int qcg_create(qcg_context** out, int device) {
    if (!out) return QCG_INVALID_ARG;
    *out = nullptr;
    auto c = std::make_unique<qcg_context>();
    c->device = device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return QCG_DEVICE;
    if (device < 0 || device >= ndev) return QCG_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return QCG_DEVICE;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return QCG_DEVICE;
    // the static scratch of qcg_build_jk_device_noalloc; a failure surfaces in that entry point
    (void)noalloc_setup(device);
    const size_t nt = sizeof(qcg_rys_coef) / sizeof(double);
    if (c->alloc(&c->d_rys, nt) != hipSuccess || c->alloc(&c->d_rys_f, nt) != hipSuccess) return QCG_DEVICE;
    std::vector<float> tf(nt);
    for (size_t i = 0; i < nt; ++i) tf[i] = (float)qcg_rys_coef[i];
    if (hipMemcpy(c->d_rys, qcg_rys_coef, nt * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return QCG_DEVICE;
    if (hipMemcpy(c->d_rys_f, tf.data(), nt * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return QCG_DEVICE;
    int cls = 0;
    for (int cb = kNumPairClasses - 1; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++cls) {
            const int la = cb == 0 ? 0 : cb <= 2 ? 1 : 2, lb = (cb == 0 || cb == 1 || cb == 3) ? 0 : (cb == 2 || cb == 4) ? 1 : 2;
            const int lc = ck == 0 ? 0 : ck <= 2 ? 1 : 2, ld = (ck == 0 || ck == 1 || ck == 3) ? 0 : (ck == 2 || ck == 4) ? 1 : 2;
            const int nq = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
            // defaults close to what the autotuner picks on alkanes: the wave-cooperative kernel
            // (block 0) above kJkLaneMaxNQ components, the per-lane kernel up to it
            for (int part = 0; part < 2; ++part) {
                c->tile[cls][part].block = nq > kJkLaneMaxNQ ? 0 : (nq <= 3 ? 256 : 64);
                c->tile[cls][part].reps = nq >= 162 ? 1 : nq > kJkLaneMaxNQ ? 2 : nq >= 18 ? 4 : 1;
            }
            c->class_nq[cls] = nq;
        }
    *out = c.release();
    return QCG_OK;
}

void qcg_destroy(qcg_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    c->drop_fock_graph();  // waits for the exec's last launch
    (void)hipStreamSynchronize(c->stream);
    if (c->ev_done) (void)hipEventSynchronize(c->ev_done);  // the last entry point's work on any stream
    for (int i = 0; i < qcg_context::kStreams; ++i)
        if (c->aux[i]) (void)hipStreamSynchronize(c->aux[i]);
    if (c->xc_stream) (void)hipStreamSynchronize(c->xc_stream);
    for (void* p : c->owned) (void)hipFree(p);
    c->owned.clear();
    if (c->ev_made)
        for (auto& e : c->ev) (void)hipEventDestroy(e);
    if (c->ev_fork) {
        (void)hipEventDestroy(c->ev_fork);
        for (int i = 0; i < qcg_context::kStreams; ++i) {
            (void)hipEventDestroy(c->ev_join[i]);
            (void)hipStreamDestroy(c->aux[i]);
        }
    }
    if (c->ev_done) (void)hipEventDestroy(c->ev_done);
    if (c->xc_stream) {
        (void)hipEventDestroy(c->ev_xc_fork);
        (void)hipEventDestroy(c->ev_xc_join);
        (void)hipStreamDestroy(c->xc_stream);
    }
    if (c->blas) (void)rocblas_destroy_handle(c->blas);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* qcg_last_error(const qcg_context* c) { return c ? c->err.c_str() : "null context"; }

int qcg_set_options(qcg_context* c, int precision, double fp32_cutoff, double schwarz_threshold) {
    if (!c) return QCG_INVALID_ARG;
    if (precision != QCG_PREC_FP64 && precision != QCG_PREC_MIXED) return fail(c, QCG_INVALID_ARG, "precision");
    if (!(fp32_cutoff > 0.0) || !(schwarz_threshold > 0.0)) return fail(c, QCG_INVALID_ARG, "thresholds must be > 0");
    c->precision = precision;
    c->fp32_cutoff = fp32_cutoff;
    c->schwarz_thr = schwarz_threshold;
    return QCG_OK;
}

int qcg_set_charge(qcg_context* c, int charge) {
    if (!c) return QCG_INVALID_ARG;
    c->charge = charge;
    return QCG_OK;
}

int qcg_nbf(const qcg_context* c, int* nbf) {
    if (!c || !nbf) return QCG_INVALID_ARG;
    if (!c->has_basis) return QCG_NO_BASIS;
    *nbf = c->basis.nbf;
    return QCG_OK;
}

int qcg_set_basis(qcg_context* c, int nshell, const int* l, const int* nprim, const int* prim_off,
                  const double* exps, const double* coefs, const double* centers, const int* shell_atom,
                  int natom, const int* Z, const double* xyz) {
    if (!c) return QCG_INVALID_ARG;
    std::vector<int> sa_local(nshell > 0 ? nshell : 0, 0);
    if (shell_atom) sa_local.assign(shell_atom, shell_atom + nshell);
    return qcg_set_basis_batch(c, 1, &nshell, &natom, l, nprim, prim_off, exps, coefs, centers, sa_local.data(), Z,
                               xyz);
}

int qcg_set_basis_batch(qcg_context* c, int nmol, const int* mol_nshell, const int* mol_natom, const int* l,
                        const int* nprim, const int* prim_off, const double* exps, const double* coefs,
                        const double* centers, const int* shell_atom, const int* Z, const double* xyz) {
    if (!c) return QCG_INVALID_ARG;
    if (nmol <= 0 || !mol_nshell || !mol_natom || !l || !nprim || !prim_off || !exps || !coefs || !centers ||
        !shell_atom || !Z || !xyz)
        return fail(c, QCG_INVALID_ARG, "qcg_set_basis: empty or null input");
    HIPCHK(c, hipSetDevice(c->device));
    c->free_basis();
    c->stats.clear();
    c->nmol = nmol;
    c->mols.assign(nmol, HostBasis());
    c->mol_nbf_h.assign(nmol, 0);
    c->mat_off_h.assign(nmol, 0);
    int s0 = 0, a0 = 0;
    int64_t moff = 0;
    c->max_nbf = 0;
    for (int m = 0; m < nmol; ++m) {
        const int ns = mol_nshell[m], na = mol_natom[m];
        if (ns <= 0 || na <= 0) return fail(c, QCG_INVALID_ARG, "molecule without shells or atoms");
        HostBasis& B = c->mols[m];
        B.nshell = ns;
        B.natom = na;
        int bf = 0;
        for (int s = s0; s < s0 + ns; ++s) {
            if (l[s] < 0 || l[s] > kMaxL) return fail(c, QCG_UNSUPPORTED_L, "angular momentum > 2 not supported");
            if (nprim[s] <= 0) return fail(c, QCG_INVALID_ARG, "shell without primitives");
            if (shell_atom[s] < 0 || shell_atom[s] >= na) return fail(c, QCG_INVALID_ARG, "shell atom out of range");
            B.l.push_back(l[s]);
            B.nprim.push_back(nprim[s]);
            B.prim_off.push_back((int)B.exps.size());
            for (int p = 0; p < nprim[s]; ++p) {
                B.exps.push_back(exps[prim_off[s] + p]);
                B.coefs.push_back(coefs[prim_off[s] + p]);
            }
            for (int x = 0; x < 3; ++x) B.centers.push_back(centers[3 * s + x]);
            B.shell_atom.push_back(shell_atom[s]);
            B.bf_off.push_back(bf);
            bf += ncart(l[s]);
        }
        B.nbf = bf;
        B.Z.assign(Z + a0, Z + a0 + na);
        B.xyz.assign(xyz + 3 * a0, xyz + 3 * (a0 + na));
        c->mol_nbf_h[m] = bf;
        c->mat_off_h[m] = moff;
        moff += (int64_t)bf * bf;
        c->max_nbf = std::max(c->max_nbf, bf);
        s0 += ns;
        a0 += na;
    }
    c->mat_total = moff;
    c->basis = c->mols[0];

    // shell pairs (si >= sj) of each molecule, relabelled so that l(sa) >= l(sb), grouped by
    // class and, inside a class, by molecule
    std::vector<std::vector<std::vector<PairRec>>> by(kNumPairClasses, std::vector<std::vector<PairRec>>(nmol));
    std::vector<PrimPair<double>> pp;
    int gs0 = 0;
    for (int m = 0; m < nmol; ++m) {
        const HostBasis& B = c->mols[m];
        for (int si = 0; si < B.nshell; ++si)
            for (int sj = 0; sj <= si; ++sj) {
                const int sa = B.l[si] >= B.l[sj] ? si : sj, sb = B.l[si] >= B.l[sj] ? sj : si;
                PairRec r{};
                const double* A = &B.centers[3 * sa];
                const double* Bc = &B.centers[3 * sb];
                r.ax = A[0]; r.ay = A[1]; r.az = A[2];
                r.abx = A[0] - Bc[0]; r.aby = A[1] - Bc[1]; r.abz = A[2] - Bc[2];
                r.sa = gs0 + sa; r.sb = gs0 + sb;
                r.bfa = B.bf_off[sa]; r.bfb = B.bf_off[sb];
                r.atom_a = B.shell_atom[sa]; r.atom_b = B.shell_atom[sb];
                r.mol = m;
                r.pp_off = (int)pp.size();
                const double ab2 = r.abx * r.abx + r.aby * r.aby + r.abz * r.abz;
                for (int i = 0; i < B.nprim[sa]; ++i)
                    for (int j = 0; j < B.nprim[sb]; ++j) {
                        const double a = B.exps[B.prim_off[sa] + i], b = B.exps[B.prim_off[sb] + j];
                        const double kab = std::exp(-(a * b / (a + b)) * ab2);
                        if (kab < kPrimScreen) continue;  // integrals/eri/mod.rs:603
                        PrimPair<double> q{};
                        q.p = a + b;
                        q.px = (a * A[0] + b * Bc[0]) / q.p;
                        q.py = (a * A[1] + b * Bc[1]) / q.p;
                        q.pz = (a * A[2] + b * Bc[2]) / q.p;
                        q.kab = kab;
                        q.ea = a;
                        q.hp = 0.5 / q.p;
                        q.ckp = kab * B.coefs[B.prim_off[sa] + i] * B.coefs[B.prim_off[sb] + j] *
                                radial_norm(a, B.l[sa]) * radial_norm(b, B.l[sb]) / q.p;
                        pp.push_back(q);
                    }
                r.npp = (int)pp.size() - r.pp_off;
                by[pair_class(B.l[sa], B.l[sb])][m].push_back(r);
            }
        gs0 += B.nshell;
    }
    std::vector<PairRec> all;
    c->ket_off_h.assign(kNumPairClasses * nmol, 0);
    c->ket_cnt_h.assign(kNumPairClasses * nmol, 0);
    for (int k = 0; k < kNumPairClasses; ++k) {
        c->class_off[k] = (int)all.size();
        for (int m = 0; m < nmol; ++m) {
            c->ket_off_h[k * nmol + m] = (int)all.size() - c->class_off[k];
            c->ket_cnt_h[k * nmol + m] = (int)by[k][m].size();
            all.insert(all.end(), by[k][m].begin(), by[k][m].end());
        }
    }
    c->class_off[kNumPairClasses] = (int)all.size();
    const int npairs = (int)all.size();
    std::vector<PrimPair<float>> ppf(pp.size());
    for (size_t i = 0; i < pp.size(); ++i) {
        ppf[i].p = (float)pp[i].p; ppf[i].px = (float)pp[i].px; ppf[i].py = (float)pp[i].py;
        ppf[i].pz = (float)pp[i].pz; ppf[i].ckp = (float)pp[i].ckp;
        ppf[i].hp = (float)pp[i].hp;
        set_prim_kab(ppf[i], pp[i].kab);  // FP64 bits in kab + ea: the FP64 screening decisions
    }
    c->h_pp = pp;
    HIPCHK(c, c->alloc(&c->d_pp, pp.size()));
    HIPCHK(c, c->alloc(&c->d_ppf, pp.size()));
    HIPCHK(c, c->alloc(&c->d_pairs, all.size()));
    HIPCHK(c, hipMemcpy(c->d_pp, pp.data(), pp.size() * sizeof(pp[0]), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ppf, ppf.data(), ppf.size() * sizeof(ppf[0]), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_pairs, all.data(), all.size() * sizeof(PairRec), hipMemcpyHostToDevice));
    // Schwarz bounds on the device
    double* dQ = nullptr;
    HIPCHK(c, c->alloc(&dQ, npairs));
    for (int k = 0; k < kNumPairClasses; ++k) {
        const int n = c->class_off[k + 1] - c->class_off[k];
        HIPCHK(c, launch_schwarz(k, c->d_pairs + c->class_off[k], n, c->d_pp, c->d_rys, dQ + c->class_off[k], c->stream));
    }
    std::vector<double> Q(npairs);
    HIPCHK(c, hipMemcpyAsync(Q.data(), dQ, npairs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->release(dQ);
    for (int i = 0; i < npairs; ++i) all[i].Q = Q[i];
    if (nmol == 1) {
        const int ns = c->basis.nshell;
        c->Qshell.assign((size_t)ns * ns, 0.0);
        for (int i = 0; i < npairs; ++i) {
            c->Qshell[(size_t)all[i].sa * ns + all[i].sb] = Q[i];
            c->Qshell[(size_t)all[i].sb * ns + all[i].sa] = Q[i];
        }
    }
    {
        // per-molecule Schwarz matrices and function -> shell maps for the fixed-point scales
        std::vector<int64_t> qoff(nmol);
        std::vector<int> mns(nmol), bfoff(nmol), shell0(nmol), sfoff(nmol), bfsh, rowmol, sfirst;
        int64_t qn = 0;
        int gs = 0;
        for (int m = 0; m < nmol; ++m) {
            const HostBasis& B = c->mols[m];
            if (B.nshell > 4096) return fail(c, QCG_INVALID_ARG, "more than 4096 shells in one molecule");
            qoff[m] = qn;
            mns[m] = B.nshell;
            shell0[m] = gs;
            bfoff[m] = (int)bfsh.size();
            sfoff[m] = (int)sfirst.size();
            for (int sh = 0; sh < B.nshell; ++sh) {
                sfirst.push_back(B.bf_off[sh]);
                for (int k = 0; k < ncart(B.l[sh]); ++k) {
                    bfsh.push_back(sh);
                    rowmol.push_back(m);
                }
            }
            sfirst.push_back(B.nbf);
            qn += (int64_t)B.nshell * B.nshell;
            gs += B.nshell;
        }
        std::vector<double> qsh((size_t)qn, 0.0), qmax(nmol, 0.0);
        for (int i = 0; i < npairs; ++i) {
            const int m = all[i].mol, ns = mns[m];
            const int a = all[i].sa - shell0[m], b = all[i].sb - shell0[m];
            qsh[(size_t)qoff[m] + (size_t)a * ns + b] = Q[i];
            qsh[(size_t)qoff[m] + (size_t)b * ns + a] = Q[i];
            qmax[m] = std::max(qmax[m], Q[i]);
        }
        HIPCHK(c, c->alloc(&c->d_fix, 4 * (size_t)nmol));
        HIPCHK(c, c->alloc(&c->d_qsh, qsh.size()));
        HIPCHK(c, c->alloc(&c->d_qsh_off, nmol));
        HIPCHK(c, c->alloc(&c->d_mol_ns, nmol));
        HIPCHK(c, c->alloc(&c->d_bf_shell, bfsh.size()));
        HIPCHK(c, c->alloc(&c->d_bf_off, nmol));
        HIPCHK(c, c->alloc(&c->d_rowsum, 2 * bfsh.size()));
        HIPCHK(c, c->alloc(&c->d_qmax, nmol));
        HIPCHK(c, c->alloc(&c->d_row_mol, rowmol.size()));
        HIPCHK(c, c->alloc(&c->d_shell_first, sfirst.size()));
        HIPCHK(c, c->alloc(&c->d_shell_first_off, nmol));
        HIPCHK(c, hipMemcpy(c->d_qsh, qsh.data(), qsh.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_qsh_off, qoff.data(), nmol * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_mol_ns, mns.data(), nmol * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_bf_shell, bfsh.data(), bfsh.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_bf_off, bfoff.data(), nmol * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_qmax, qmax.data(), nmol * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_row_mol, rowmol.data(), rowmol.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_shell_first, sfirst.data(), sfirst.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_shell_first_off, sfoff.data(), nmol * sizeof(int), hipMemcpyHostToDevice));
        c->fixdata.qsh = c->d_qsh;
        c->fixdata.qsh_off = c->d_qsh_off;
        c->fixdata.mol_ns = c->d_mol_ns;
        c->fixdata.bf_shell = c->d_bf_shell;
        c->fixdata.bf_off = c->d_bf_off;
        c->fixdata.row_mol = c->d_row_mol;
        c->fixdata.nrows = (int)rowmol.size();
        c->fixdata.rowsum = c->d_rowsum;
        c->fixdata.qmax = c->d_qmax;
        c->fixdata.shell_first = c->d_shell_first;
        c->fixdata.shell_first_off = c->d_shell_first_off;
    }
    // contraction buckets: every wavefront of a J/K launch then runs the same primitive loop
    // trip count; Schwarz order inside each bucket keeps the screened kets a prefix
    c->bucket_bnd.assign(kNumPairClasses * nmol, std::vector<int>());
    for (int k = 0; k < kNumPairClasses; ++k)
        for (int m = 0; m < nmol; ++m) {
            auto first = all.begin() + c->class_off[k] + c->ket_off_h[k * nmol + m];
            auto last = first + c->ket_cnt_h[k * nmol + m];
            std::stable_sort(first, last, [](const PairRec& x, const PairRec& y) {
                return x.npp != y.npp ? x.npp > y.npp : x.Q > y.Q;
            });
            std::vector<int>& bnd = c->bucket_bnd[k * nmol + m];
            bnd.push_back(0);
            int pos = 0;
            for (auto it = first; it != last; ++it) {
                if (pos > 0 && it->npp != (it - 1)->npp) bnd.push_back(pos);
                it->pos = pos++;
            }
            if (pos > 0) bnd.push_back(pos);
        }
    HIPCHK(c, hipMemcpy(c->d_pairs, all.data(), all.size() * sizeof(PairRec), hipMemcpyHostToDevice));
    c->pairs = all;
    HIPCHK(c, c->alloc(&c->d_mol_nbf, nmol));
    HIPCHK(c, c->alloc(&c->d_mat_off, nmol));
    HIPCHK(c, c->alloc(&c->d_ket_off, kNumPairClasses * nmol));
    HIPCHK(c, c->alloc(&c->d_ket_cnt, kNumPairClasses * nmol));
    HIPCHK(c, hipMemcpy(c->d_mol_nbf, c->mol_nbf_h.data(), nmol * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_mat_off, c->mat_off_h.data(), nmol * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ket_off, c->ket_off_h.data(), c->ket_off_h.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ket_cnt, c->ket_cnt_h.data(), c->ket_cnt_h.size() * sizeof(int), hipMemcpyHostToDevice));
    const size_t n2 = (size_t)c->mat_total;
    HIPCHK(c, c->alloc(&c->d_D, n2));
    HIPCHK(c, c->alloc(&c->d_J, n2));
    HIPCHK(c, c->alloc(&c->d_K, n2));
    HIPCHK(c, c->alloc(&c->d_Jacc, n2));
    HIPCHK(c, c->alloc(&c->d_Kacc, n2));
    HIPCHK(c, c->alloc(&c->d_H, n2));
    HIPCHK(c, c->alloc(&c->d_F, n2));
    c->has_basis = true;
    return QCG_OK;
}

int qcg_schwarz_bounds(qcg_context* c, double* Q) {
    if (!c || !Q) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    std::memcpy(Q, c->Qshell.data(), c->Qshell.size() * sizeof(double));
    return QCG_OK;
}

int qcg_screen_counts(qcg_context* c, int64_t* n_kept, int64_t* n_fp32) {
    if (!c) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    int64_t kept = 0, f32 = 0;
    const double cut = c->fp32_cutoff > c->schwarz_thr ? c->fp32_cutoff : c->schwarz_thr;
    for (int cb = 0; cb < kNumPairClasses; ++cb)
        for (int ck = 0; ck <= cb; ++ck)
            for (int m = 0; m < c->nmol; ++m)
                for (int p = 0; p < c->ket_cnt_h[cb * c->nmol + m]; ++p)
                    for_ket_ranges(c, cb, ck, m, p, c->schwarz_thr, cut, [&](int bs, int sp, int e) {
                        kept += e - bs;
                        f32 += e - sp;
                    });
    if (n_kept) *n_kept = kept;
    if (n_fp32) *n_fp32 = f32;
    return QCG_OK;
}

}  // extern "C"

static bool plan_fresh(const qcg_context* c) {
    const double key[4] = {(double)c->precision, c->fp32_cutoff, c->schwarz_thr, (double)c->tile_sig};
    return c->plan_sig == 1 && std::memcmp(key, c->plan_key, sizeof(key)) == 0 && c->ev_fork != nullptr &&
           c->ev_done != nullptr;
}

bool qcg::jk_prepared(const qcg_context* c) { return plan_fresh(c) && !(c->profiling && !c->ev_made); }

static int build_plan(qcg_context* c) {
    const bool mixed = c->precision == QCG_PREC_MIXED && c->fp32_cutoff > c->schwarz_thr;
    const double key[4] = {(double)c->precision, c->fp32_cutoff, c->schwarz_thr, (double)c->tile_sig};
    if (c->plan_sig == 1 && std::memcmp(key, c->plan_key, sizeof(key)) == 0) return QCG_OK;
    // a work item encodes its bra's primitive range in 8 bits (cut_items_split)
    for (const PairRec& pr : c->pairs)
        if (pr.npp > kItemMaxPrim)
            return fail(c, QCG_INVALID_ARG, "shell pair with more than 255 primitive pairs (at most 15 primitives per shell)");
    for (int cb = 0; cb < kNumPairClasses; ++cb) {
        int m = 1;
        for (int i = c->class_off[cb]; i < c->class_off[cb + 1]; ++i) m = std::max(m, c->pairs[i].npp);
        const int split = (c->prim_split > 0 && c->prim_split < kBraPPMax) ? c->prim_split : kBraPPMax;
        c->bra_pp[cb] = std::min(m, split);
    }
    c->items.clear();
    c->segs.clear();
    const int nmol = c->nmol;
    const double cut = mixed ? c->fp32_cutoff : c->schwarz_thr;
    int cls = 0;
    for (int cb = kNumPairClasses - 1; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++cls) {
            // the class's model flops over its surviving quartets, before the primitive screen
            // (the default launch schedule's cost): n_roots [80 + 15 (la+lb+1)(lc+ld+1)
            // + N_q (4 + 6 (lb+ld))] per primitive quartet plus 12 N_q per quartet
            const int la = cb == 0 ? 0 : cb <= 2 ? 1 : 2, lb = (cb == 0 || cb == 1 || cb == 3) ? 0 : (cb == 2 || cb == 4) ? 1 : 2;
            const int lc = ck == 0 ? 0 : ck <= 2 ? 1 : 2, ld = (ck == 0 || ck == 1 || ck == 3) ? 0 : (ck == 2 || ck == 4) ? 1 : 2;
            const double nq = (double)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
            const double fpq = ((la + lb + lc + ld) / 2 + 1) * (80.0 + 15.0 * (la + lb + 1) * (lc + ld + 1) + nq * (4.0 + 6.0 * (lb + ld)));
            double cost = 0.0;
            for (int part = 0; part < 2; ++part) {
                // fused mixed launches run both parts with the FP64 part's tile
                const qcg_context::Tile& t = c->tile[cls][(mixed && c->fuse_cls[cls]) ? 0 : part];
                const int ch = std::min(kItemMaxKets, std::max(1, t.block > 0 ? t.block * t.reps : t.reps));
                c->item_off[cls][part] = (int)c->items.size();
                if (part == 1 && !mixed) {
                    c->item_cnt[cls][part] = 0;
                    continue;
                }
                for (int m = 0; m < nmol; ++m) {
                    const int pbase = c->ket_off_h[cb * nmol + m];
                    for (int p = 0; p < c->ket_cnt_h[cb * nmol + m]; ++p) {
                        const int g0 = (int)c->segs.size();
                        const int npp = c->pairs[c->class_off[cb] + pbase + p].npp;
                        const int kbase = c->class_off[ck] + c->ket_off_h[ck * nmol + m];
                        for_ket_ranges(c, cb, ck, m, p, c->schwarz_thr, cut, [&](int bs, int sp, int e) {
                            const int lo = part == 0 ? bs : sp, hi = part == 0 ? sp : e;
                            if (hi > lo) {
                                c->segs.push_back(make_int2(lo, hi - lo));
                                // a segment lies in one contraction bucket: one ket primitive count
                                cost += (hi - lo) * (fpq * npp * c->pairs[kbase + lo].npp + 12.0 * nq);
                            }
                        });
                        // bras above kBraPPMax primitive pairs are always cut into ranges
                        const int split = (c->prim_split > 0 && c->prim_split < kBraPPMax) ? c->prim_split : kBraPPMax;
                        cut_items_split(c->items, c->segs, g0, pbase + p, ch, npp, split);
                    }
                }
                c->item_cnt[cls][part] = (int)c->items.size() - c->item_off[cls][part];
            }
            c->plan_cost[cls] = cost;
        }
    c->release(c->d_items);
    c->d_items = nullptr;
    HIPCHK(c, c->alloc(&c->d_items, c->items.size()));
    HIPCHK(c, hipMemcpy(c->d_items, c->items.data(), c->items.size() * sizeof(int4), hipMemcpyHostToDevice));
    c->release(c->d_segs);
    c->d_segs = nullptr;
    HIPCHK(c, c->alloc(&c->d_segs, c->segs.size()));
    HIPCHK(c, hipMemcpy(c->d_segs, c->segs.data(), c->segs.size() * sizeof(int2), hipMemcpyHostToDevice));
    std::memcpy(c->plan_key, key, sizeof(key));
    c->plan_sig = 1;
    return QCG_OK;
}

int qcg::prepare_jk(qcg_context* c) {
    int rc = build_plan(c);
    if (rc) return rc;
    if (!c->ev_done) HIPCHK(c, hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming));
    if (!c->ev_fork) {
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        for (int i = 0; i < qcg_context::kStreams; ++i) {
            HIPCHK(c, hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking));
            HIPCHK(c, hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming));
        }
    }
    return QCG_OK;
}

int qcg::run_jk(qcg_context* c, const double* dD, double* dJ, double* dK, hipStream_t st, const JkAcc* acc,
                bool no_alloc, bool symmetrize) {
    const int nbf = c->max_nbf;
    const int do_j = dJ != nullptr, do_k = dK != nullptr;
    if (no_alloc || capturing(st)) {
        // a stale plan would free, allocate and copy synchronously, which a stream capture does not
        // allow (it would end the capture with an error) and the no-allocation entry point must not
        // do: those callers need the plan prepared beforehand
        if (!jk_prepared(c))
            return fail(c, QCG_INVALID_ARG,
                        "run_jk: work plan not prepared (build once with qcg_build_jk_device first, outside any "
                        "stream capture, after the last change of options or tiles)");
    } else {
        int rc = prepare_jk(c);
        if (rc) return rc;
    }
    const JkAcc own{c->d_Jacc, c->d_Kacc, c->d_fix, c->d_rowsum};
    const JkAcc& A = acc ? *acc : own;
    HIPCHK(c, stream_begin(c, st));
    // the fixed-point scales come from a bound at this density; the same pass zeroes the
    // accumulators (a kernel, no memset: graph replays stay exact)
    FixData fd = c->fixdata;
    fd.mat_off = c->d_mat_off;
    fd.mol_nbf = c->d_mol_nbf;
    fd.rowsum = A.rowsum;
    HIPCHK(c, launch_fix_scale(dD, fd, A.fix, c->nmol, do_j ? A.Jacc : nullptr, do_k ? A.Kacc : nullptr, st));
    const bool mixed = c->precision == QCG_PREC_MIXED && c->fp32_cutoff > c->schwarz_thr;
    if (c->profiling && !c->ev_made) {
        for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
        c->ev_made = true;
    }
    // class kernels spread over the auxiliary streams (forked from and joined back to st);
    // in profiling mode everything runs on st so the events bracket single kernels
    const bool fork = !c->profiling;
    if (fork) {
        HIPCHK(c, hipEventRecord(c->ev_fork, st));
        for (int i = 0; i < qcg_context::kStreams; ++i) HIPCHK(c, hipStreamWaitEvent(c->aux[i], c->ev_fork, 0));
    }
    std::memset(c->ev_recorded, 0, sizeof(c->ev_recorded));
    // class index -> (bra class, ket class), in the order (dd|dd), (dd|dp), ..., (ss|ss)
    int cls_cb[21], cls_ck[21];
    for (int cb = kNumPairClasses - 1, i = 0; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++i) {
            cls_cb[i] = cb;
            cls_ck[i] = ck;
        }
    {
        // longest processing time first onto the least-loaded stream, by the measured or set
        // costs, else by the plan's model flops
        const double* cost = c->class_cost_set ? c->class_ms : c->plan_cost;
        int* ord = c->launch_order;
        for (int i = 0; i < 21; ++i) ord[i] = i;
        std::stable_sort(ord, ord + 21, [&](int x, int y) { return cost[x] > cost[y]; });
        double load[qcg_context::kStreams] = {};
        for (int k = 0; k < 21; ++k) {
            const int s = (int)(std::min_element(load, load + qcg_context::kStreams) - load);
            c->launch_stream[ord[k]] = s;
            load[s] += cost[ord[k]];
        }
    }
    for (int k = 0; k < 21; ++k) {
        const int cls = c->launch_order[k];
        const int cb = cls_cb[cls], ck = cls_ck[cls];
        for (int part = 0; part < 2; ++part) {
            if (part == 1 && !mixed) continue;
            // mixed precision: both parts of the class in one launch (the plan cut the FP32
            // items with the FP64 part's tile and put them right after the FP64 items)
            if (part == 1 && c->fuse_cls[cls]) continue;
            const bool fused = mixed && c->fuse_cls[cls];
            JkLaunch a{};
            a.bra = c->d_pairs + c->class_off[cb];
            a.ket = c->d_pairs + c->class_off[ck];
            a.diag = cb == ck;
            a.D = dD;
            a.nbf = nbf;
            a.Jacc = A.Jacc;
            a.Kacc = A.Kacc;
            a.fix = A.fix;
            a.do_j = do_j;
            a.do_k = do_k;
            a.ket_off = c->d_ket_off + ck * c->nmol;
            a.ket_cnt = c->d_ket_cnt + ck * c->nmol;
            a.mat_off = c->d_mat_off;
            a.mol_nbf = c->d_mol_nbf;
            a.items = c->d_items + c->item_off[cls][part];
            a.segs = c->d_segs;
            a.nitems = c->item_cnt[cls][part];
            a.block = c->tile[cls][part].block;
            a.bra_pp = c->bra_pp[cb];
            a.pp = part ? (const void*)c->d_ppf : (const void*)c->d_pp;
            a.rys = part ? (const void*)c->d_rys_f : (const void*)c->d_rys;
            if (fused) {
                a.nitems = c->item_cnt[cls][0] + c->item_cnt[cls][1];
                a.n64 = c->item_cnt[cls][0];
                a.ppf = c->d_ppf;
                a.rysf = c->d_rys_f;
            }
            hipStream_t s = !fork ? st : c->aux[c->launch_stream[cls]];
            if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[4 * cls + 2 * part], s));
            HIPCHK(c, launch_jk(cb, ck, part == 1, a, s));
            if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[4 * cls + 2 * part + 1], s));
            c->ev_recorded[cls][part] = c->profiling;
        }
    }
    if (fork)
        for (int i = 0; i < qcg_context::kStreams; ++i) {
            HIPCHK(c, hipEventRecord(c->ev_join[i], c->aux[i]));
            HIPCHK(c, hipStreamWaitEvent(st, c->ev_join[i], 0));
        }
    if (!symmetrize) {
        // the caller reads the accumulators itself (launch_fock_combine)
    } else if (c->nmol == 1) HIPCHK(c, launch_symmetrize(A.Jacc, A.Kacc, A.fix, dJ, dK, nbf, st));
    else
        HIPCHK(c, launch_symmetrize_batch(A.Jacc, A.Kacc, A.fix, dJ, dK, c->nmol, c->d_mat_off, c->d_mol_nbf, nbf,
                                          st));
    HIPCHK(c, stream_done(c, st));
    return QCG_OK;
}

extern "C" {

}  // extern "C"

namespace {

// Autotuner candidates {threads per block, kets per thread}; block 0 = one wave per work item of
// `reps` kets (wave-cooperative kernel, classes with at least 27 components only).
constexpr int kTuneCand[][2] = {{64, 1}, {64, 2}, {64, 4}, {128, 1}, {256, 1}, {64, 8},
                                {128, 4}, {256, 2}, {0, 1}, {0, 2}, {0, 4}};
constexpr int kTuneNCand = sizeof(kTuneCand) / sizeof(kTuneCand[0]);

// Candidate k mapped onto what class cls has: classes with fewer than 27 components have only the
// per-lane kernel, those above kJkLaneMaxNQ only the wave-cooperative one.
qcg_context::Tile tune_tile(const qcg_context* c, int k, int cls) {
    const int nq = c->class_nq[cls];
    if (nq < 27) return kTuneCand[k][0] == 0 ? qcg_context::Tile{64, 1} : qcg_context::Tile{kTuneCand[k][0], kTuneCand[k][1]};
    if (nq > kJkLaneMaxNQ)
        return kTuneCand[k][0] == 0 ? qcg_context::Tile{0, kTuneCand[k][1]}
                                    : qcg_context::Tile{0, std::min(8, kTuneCand[k][1] * 2)};
    return qcg_context::Tile{kTuneCand[k][0], kTuneCand[k][1]};
}

void export_tiles(const qcg_context* c, int* chosen) {
    if (!chosen) return;
    for (int i = 0; i < 21; ++i)
        for (int part = 0; part < 2; ++part) {
            chosen[4 * i + 2 * part] = c->tile[i][part].block;
            chosen[4 * i + 2 * part + 1] = c->tile[i][part].reps;
        }
}

}  // namespace

extern "C" {

// Autotuner: time every class kernel under each candidate tile (threads per block x kets per
// thread) on the given density and keep the fastest candidate per class.
int qcg_autotune(qcg_context* c, const double* dD, int iters, int* chosen) {
    if (!c || !dD) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    const int ncand = kTuneNCand;
    auto tile_of = [&](int k, int cls) { return tune_tile(c, k, cls); };
    for (int i = 0; i < 21; ++i) c->fuse_cls[i] = true;  // the per-class timings run fused launches
    std::vector<double> best(42, 1e300);
    std::vector<int> pick(42, 0);
    const bool prof = c->profiling;
    c->profiling = true;
    iters = std::max(1, iters);
    for (int k = 0; k < ncand; ++k) {
        for (int i = 0; i < 21; ++i) c->tile[i][0] = c->tile[i][1] = tile_of(k, i);
        c->tile_sig += 1;
        std::vector<double> tot(42, 0.0);
        for (int it = 0; it < iters + 1; ++it) {
            int rc = run_jk(c, dD, c->d_J, c->d_K, c->stream);
            if (rc) { c->profiling = prof; return rc; }
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (it == 0) continue;  // warm-up
            for (int i = 0; i < 21; ++i)
                for (int part = 0; part < 2; ++part)
                    if (c->ev_recorded[i][part]) {
                        float ms = 0.f;
                        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[4 * i + 2 * part], c->ev[4 * i + 2 * part + 1]));
                        tot[2 * i + part] += ms;
                    }
        }
        for (int i = 0; i < 42; ++i)
            if (tot[i] < best[i]) { best[i] = tot[i]; pick[i] = k; }
    }
    for (int i = 0; i < 21; ++i) c->class_ms[i] = (best[2 * i] + best[2 * i + 1]) / iters;
    c->class_cost_set = true;
    for (int i = 0; i < 21; ++i)
        for (int part = 0; part < 2; ++part) c->tile[i][part] = tile_of(pick[2 * i + part], i);
    export_tiles(c, chosen);
    c->tile_sig += 1;
    c->profiling = prof;
    return QCG_OK;
}

// Second tuning pass on the whole build: the per-class timings above run one class at a time,
// while a build runs the 21 class kernels on four streams at once, where a tile's register and
// LDS footprint also decides how the kernels share the GPU.  Greedy over the classes (longest
// first) and the candidates of each: keep a change of the FP64-part tile (the tile the fused
// mixed launch cuts both parts with) when `builds` whole builds get at least 0.5 % faster, twice.
int qcg_autotune_refine(qcg_context* c, const double* dD, int builds, int* chosen) {
    if (!c || !dD) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "autotune_refine: one molecule per context");
    builds = std::max(builds, 4);
    int rc = prepare_jk(c);
    if (rc) return rc;
    if (!c->ev_made) {  // the timing events (made by the first profiled build otherwise)
        for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
        c->ev_made = true;
    }
    const bool prof = c->profiling;
    c->profiling = false;
    double ms = 0.0;
    // Timed as the device Fock builds run: qcg_build_fock_device replaying its captured graph on
    // a stream of its own (recaptured for every tile change), with a zero H into a scratch F.
    // The work plan is rebuilt before each capture (qcg_build_fock_device calls prepare_jk before
    // hipStreamBeginCapture); run_jk refuses to rebuild it inside a capture.
    const size_t n2 = (size_t)c->basis.nbf * c->basis.nbf;
    double *dH = nullptr, *dF = nullptr;
    hipStream_t ts = nullptr;
    const bool graph = c->use_graph;
    auto cleanup = [&]() {
        c->drop_fock_graph();  // waits for its launches on ts, before ts goes away
        if (ts) {
            (void)hipStreamSynchronize(ts);
            if (c->last_stream == ts) c->last_stream = nullptr;  // drained: nothing to order after
            (void)hipStreamDestroy(ts);
        }
        c->release(dH);
        c->release(dF);
        c->use_graph = graph;
        c->profiling = prof;
    };
    if (c->alloc(&dH, n2) != hipSuccess || c->alloc(&dF, n2) != hipSuccess ||
        hipStreamCreateWithFlags(&ts, hipStreamNonBlocking) != hipSuccess ||
        launch_fill(dH, n2, 0.0, ts) != hipSuccess) {
        cleanup();
        return fail(c, QCG_DEVICE, "autotune_refine: set-up");
    }
    c->use_graph = true;
    auto measure = [&]() -> int {
        for (int w = 0; w < 2; ++w) {  // warm-up; the first also captures the graph for new tiles
            int r = qcg_build_fock_device(c, dH, dD, 0.5, dF, ts);
            if (r) return r;
        }
        HIPCHK(c, hipEventRecord(c->ev[84], ts));
        for (int b = 0; b < builds; ++b) {
            int r = qcg_build_fock_device(c, dH, dD, 0.5, dF, ts);
            if (r) return r;
        }
        HIPCHK(c, hipEventRecord(c->ev[85], ts));
        HIPCHK(c, hipEventSynchronize(c->ev[85]));
        float el = 0.f;
        HIPCHK(c, hipEventElapsedTime(&el, c->ev[84], c->ev[85]));
        ms = (double)el / builds;
        return QCG_OK;
    };
    double best = 1e300;
    for (int r = 0; r < 2; ++r) {
        if ((rc = measure())) break;
        best = std::min(best, ms);
    }
    int order[21];
    for (int i = 0; i < 21; ++i) order[i] = i;
    std::stable_sort(order, order + 21, [&](int x, int y) { return c->class_ms[x] > c->class_ms[y]; });
    for (int oi = 0; oi < 21 && !rc; ++oi) {
        const int cls = order[oi];
        if (c->item_cnt[cls][0] + c->item_cnt[cls][1] == 0) continue;
        for (int k = 0; k < kTuneNCand && !rc; ++k) {
            const qcg_context::Tile cur = c->tile[cls][0], t = tune_tile(c, k, cls);
            if (t.block == cur.block && t.reps == cur.reps) continue;
            c->tile[cls][0] = t;
            c->tile_sig += 1;
            if ((rc = measure())) break;
            double got = ms;
            bool keep = got < 0.995 * best;
            if (keep) {
                if ((rc = measure())) break;
                keep = ms < 0.995 * best;
                got = std::min(got, ms);
            }
            if (keep) {
                best = got;
            } else {
                c->tile[cls][0] = cur;
                c->tile_sig += 1;
            }
        }
        // mixed precision: the class's FP32 items in a launch of their own, with the FP32 part's
        // tile, instead of the fused launch at the FP64 part's register count
        if (!rc && c->precision == QCG_PREC_MIXED && c->item_cnt[cls][1] > 0) {
            c->fuse_cls[cls] = false;
            c->tile_sig += 1;
            if (!(rc = measure())) {
                double got = ms;
                bool keep = got < 0.995 * best;
                if (keep && !(rc = measure())) {
                    keep = ms < 0.995 * best;
                    got = std::min(got, ms);
                }
                if (!rc && keep) best = got;
                else {
                    c->fuse_cls[cls] = true;
                    c->tile_sig += 1;
                }
            }
        }
    }
    cleanup();
    if (rc) return rc;
    export_tiles(c, chosen);
    return QCG_OK;
}

int qcg_set_tiles(qcg_context* c, const int* tiles) {
    if (!c || !tiles) return QCG_INVALID_ARG;
    for (int i = 0; i < 42; ++i)
        if (tiles[2 * i] < 0 || tiles[2 * i] > 256 || (tiles[2 * i] % 64) != 0 || tiles[2 * i + 1] < 1 ||
            tiles[2 * i + 1] > 64)
            return fail(c, QCG_INVALID_ARG, "tile: block must be 0 or a multiple of 64 up to 256, kets 1..64");
    for (int i = 0; i < 21; ++i)
        for (int part = 0; part < 2; ++part)
            c->tile[i][part] = {tiles[4 * i + 2 * part], tiles[4 * i + 2 * part + 1]};
    c->tile_sig += 1;
    return QCG_OK;
}

int qcg_get_tiles(qcg_context* c, int* tiles) {
    if (!c || !tiles) return QCG_INVALID_ARG;
    for (int i = 0; i < 21; ++i)
        for (int part = 0; part < 2; ++part) {
            tiles[4 * i + 2 * part] = c->tile[i][part].block;
            tiles[4 * i + 2 * part + 1] = c->tile[i][part].reps;
        }
    return QCG_OK;
}

int qcg_set_profiling(qcg_context* c, int on) {
    if (!c) return QCG_INVALID_ARG;
    c->profiling = on != 0;
    return QCG_OK;
}

int qcg_class_stats(qcg_context* c, qcg_class_stat* out, int* n) {
    if (!c || !out) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    const bool mixed = c->precision == QCG_PREC_MIXED && c->fp32_cutoff > c->schwarz_thr;
    const double thr = c->schwarz_thr, cut = mixed ? c->fp32_cutoff : thr;
    const bool cached = c->stats.size() == 21 && c->stats_key[0] == (double)c->precision &&
                        c->stats_key[1] == c->fp32_cutoff && c->stats_key[2] == c->schwarz_thr;
    if (cached) {
        for (int i = 0; i < 21; ++i) out[i] = c->stats[i];
    }
    int cls = 0;
    for (int cb = kNumPairClasses - 1; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++cls) {
            qcg_class_stat& s = out[cls];
            if (cached) {
                s.ms_fp64 = s.ms_fp32 = 0.0;
            } else {
            std::memset(&s, 0, sizeof(s));
            const int la = cb == 0 ? 0 : cb <= 2 ? 1 : 2, lb = (cb == 0 || cb == 1 || cb == 3) ? 0 : (cb == 2 || cb == 4) ? 1 : 2;
            const int lc = ck == 0 ? 0 : ck <= 2 ? 1 : 2, ld = (ck == 0 || ck == 1 || ck == 3) ? 0 : (ck == 2 || ck == 4) ? 1 : 2;
            s.la = la; s.lb = lb; s.lc = lc; s.ld = ld;
            int64_t q64 = 0, q32 = 0, p64 = 0, p32 = 0;
            for (int m = 0; m < c->nmol; ++m) {
                const int nb = c->ket_cnt_h[cb * c->nmol + m];
                const int bbase = c->class_off[cb] + c->ket_off_h[cb * c->nmol + m];
                const int kbase = c->class_off[ck] + c->ket_off_h[ck * c->nmol + m];
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : q64, q32, p64, p32)
                for (int p = 0; p < nb; ++p) {
                    const PairRec& P = c->pairs[bbase + p];
                    for_ket_ranges(c, cb, ck, m, p, thr, cut, [&](int bs, int sp, int e) {
                        for (int q = bs; q < e; ++q) {
                            const PairRec& Q = c->pairs[kbase + q];
                            int64_t np = 0;
                            for (int i = 0; i < P.npp; ++i) {
                                const double ka = c->h_pp[P.pp_off + i].kab;
                                for (int j = 0; j < Q.npp; ++j) np += (ka * c->h_pp[Q.pp_off + j].kab >= kPrimScreen) ? 1 : 0;
                            }
                            if (q < sp) { ++q64; p64 += np; } else { ++q32; p32 += np; }
                        }
                    });
                }
            }
            const int nr = (la + lb + lc + ld) / 2 + 1;
            const double nq = (double)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
            const double fpq = nr * (80.0 + 15.0 * (la + lb + 1) * (lc + ld + 1) + nq * (4.0 + 6.0 * (lb + ld)));
            s.quartets_fp64 = q64; s.quartets_fp32 = q32;
            s.prim_quartets_fp64 = p64; s.prim_quartets_fp32 = p32;
            s.flops_fp64 = fpq * (double)p64 + 12.0 * nq * (double)q64;
            s.flops_fp32 = fpq * (double)p32 + 12.0 * nq * (double)q32;
            }
            if (c->ev_made) {
                float ms = 0.f;
                if (c->ev_recorded[cls][0] && hipEventSynchronize(c->ev[4 * cls + 1]) == hipSuccess &&
                    hipEventElapsedTime(&ms, c->ev[4 * cls + 0], c->ev[4 * cls + 1]) == hipSuccess)
                    s.ms_fp64 = ms;
                if (c->ev_recorded[cls][1] && hipEventSynchronize(c->ev[4 * cls + 3]) == hipSuccess &&
                    hipEventElapsedTime(&ms, c->ev[4 * cls + 2], c->ev[4 * cls + 3]) == hipSuccess)
                    s.ms_fp32 = ms;
            }
        }
    if (n) *n = cls;
    return QCG_OK;
}

int qcg_build_jk_device(qcg_context* c, const double* dD, double* dJ, double* dK, void* stream) {
    if (!c || !dD) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    return run_jk(c, dD, dJ, dK, stream ? (hipStream_t)stream : c->stream);
}

int qcg_build_fock_device(qcg_context* c, const double* dH, const double* dD, double k_factor, double* dF,
                          void* stream) {
    if (!c || !dH || !dD || !dF) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (c->use_graph && !c->profiling) {
        // replay the captured launch sequence; capture again when anything it bakes in changed
        const double key[10] = {(double)(uintptr_t)dH, (double)(uintptr_t)dD, (double)(uintptr_t)dF,
                                (double)(uintptr_t)st, k_factor, (double)c->tile_sig, (double)c->precision,
                                c->fp32_cutoff, c->schwarz_thr, 1.0};
        if (!c->fock_exec || std::memcmp(key, c->fock_key, sizeof(key)) != 0) {
            c->drop_fock_graph();
            int rc = prepare_jk(c);
            if (rc) return rc;
            HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = run_jk(c, dD, c->d_J, c->d_K, st, nullptr, false, false);
            hipError_t e = rc ? hipSuccess : launch_fock_combine(dH, c->d_Jacc, c->d_Kacc, c->d_fix, dF, c->basis.nbf, k_factor, st);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc || e != hipSuccess || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                return rc ? rc : fail(c, QCG_DEVICE, std::string("Fock graph capture: ") +
                                                         hipGetErrorString(e != hipSuccess ? e : ec));
            }
            c->fock_graph = g;
            HIPCHK(c, hipGraphInstantiate(&c->fock_exec, g, nullptr, nullptr, 0));
            std::memcpy(c->fock_key, key, sizeof(key));
        }
        HIPCHK(c, stream_begin(c, st));
        HIPCHK(c, hipGraphLaunch(c->fock_exec, st));
        HIPCHK(c, stream_done(c, st));
        return QCG_OK;
    }
    int rc = run_jk(c, dD, c->d_J, c->d_K, st, nullptr, false, false);
    if (rc) return rc;
    HIPCHK(c, launch_fock_combine(dH, c->d_Jacc, c->d_Kacc, c->d_fix, dF, c->basis.nbf, k_factor, st));
    HIPCHK(c, stream_done(c, st));
    return QCG_OK;
}

int qcg_set_prim_split(qcg_context* c, int split) {
    if (!c) return QCG_INVALID_ARG;
    if (split < 0) return fail(c, QCG_INVALID_ARG, "prim split must be >= 0");
    c->prim_split = split;
    ++c->tile_sig;  // the work items change: plan and captured graph are rebuilt
    return QCG_OK;
}

int qcg_set_class_cost(qcg_context* c, const double* cost) {
    if (!c) return QCG_INVALID_ARG;
    bool any = false;
    if (cost)
        for (int i = 0; i < 21; ++i) {
            if (!(cost[i] >= 0.0) || !std::isfinite(cost[i]))
                return fail(c, QCG_INVALID_ARG, "class cost must be finite and >= 0");
            any = any || cost[i] > 0.0;
        }
    c->class_cost_set = any;
    for (int i = 0; i < 21; ++i) c->class_ms[i] = any ? cost[i] : 0.0;
    ++c->tile_sig;  // the launch schedule changes: the captured graph is rebuilt
    return QCG_OK;
}

int qcg_get_class_cost(qcg_context* c, double* cost, int* measured) {
    if (!c || !cost) return QCG_INVALID_ARG;
    for (int i = 0; i < 21; ++i) cost[i] = c->class_cost_set ? c->class_ms[i] : c->plan_cost[i];
    if (measured) *measured = c->class_cost_set ? 1 : 0;
    return QCG_OK;
}

int qcg_set_graph(qcg_context* c, int on) {
    if (!c) return QCG_INVALID_ARG;
    c->use_graph = on != 0;
    if (!c->use_graph) c->drop_fock_graph();
    return QCG_OK;
}

int qcg_build_jk(qcg_context* c, const double* D, double* J, double* K) {
    if (!c || !D) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n2 = (size_t)c->basis.nbf * c->basis.nbf;
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = run_jk(c, c->d_D, J ? c->d_J : nullptr, K ? c->d_K : nullptr, c->stream);
    if (rc) return rc;
    if (J) HIPCHK(c, hipMemcpyAsync(J, c->d_J, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (K) HIPCHK(c, hipMemcpyAsync(K, c->d_K, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QCG_OK;
}

int qcg_build_fock(qcg_context* c, const double* H, const double* D, double* F) {
    if (!c || !H || !D || !F) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    const int nbf = c->basis.nbf;
    const size_t n2 = (size_t)nbf * nbf;
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_H, H, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = run_jk(c, c->d_D, c->d_J, c->d_K, c->stream, nullptr, false, false);
    if (rc) return rc;
    HIPCHK(c, launch_fock_combine(c->d_H, c->d_Jacc, c->d_Kacc, c->d_fix, c->d_F, nbf, 0.5, c->stream));
    HIPCHK(c, hipMemcpyAsync(F, c->d_F, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QCG_OK;
}

int qcg_eri_compressed(qcg_context* c, double* out) {
    if (!c || !out) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nbf = c->basis.nbf;
    const int64_t npair = nbf * (nbf + 1) / 2;
    const size_t neri = (size_t)(npair * (npair + 1) / 2);
    double* dout = nullptr;
    HIPCHK(c, c->alloc(&dout, neri));
    HIPCHK(c, launch_fill(dout, neri, 0.0, c->stream));
    for (int cb = 0; cb < kNumPairClasses; ++cb)
        for (int ck = 0; ck <= cb; ++ck) {
            JkLaunch a{};
            a.bra = c->d_pairs + c->class_off[cb];
            a.nbra = c->class_off[cb + 1] - c->class_off[cb];
            a.ket = c->d_pairs + c->class_off[ck];
            a.nket = c->class_off[ck + 1] - c->class_off[ck];
            a.diag = cb == ck;
            a.pp = c->d_pp;
            a.rys = c->d_rys;
            a.thr_lo = c->schwarz_thr;
            HIPCHK(c, launch_eri_tensor(cb, ck, a, dout, c->stream));
        }
    HIPCHK(c, hipMemcpyAsync(out, dout, neri * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->release(dout);
    return QCG_OK;
}

int qcg_boys_eval_all(qcg_context* c, int m_max, int n, const double* T, double* F) {
    if (!c || (n > 0 && (!T || !F)) || n < 0) return QCG_INVALID_ARG;
    if (m_max < 0 || m_max > 50) return fail(c, QCG_INVALID_ARG, "BoysError::InvalidOrder");
    for (int i = 0; i < n; ++i)
        if (!(T[i] >= 0.0)) return fail(c, QCG_INVALID_ARG, "BoysError::NegativeArgument");
    if (n == 0) return QCG_OK;
    HIPCHK(c, hipSetDevice(c->device));
    double *dT = nullptr, *dF = nullptr;
    int* dst = nullptr;
    HIPCHK(c, c->alloc(&dT, n));
    HIPCHK(c, c->alloc(&dF, (size_t)n * (m_max + 1)));
    HIPCHK(c, c->alloc(&dst, 1));
    HIPCHK(c, hipMemcpyAsync(dT, T, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(dst, 0, sizeof(int), c->stream));
    HIPCHK(c, launch_boys_batch(m_max, n, dT, dF, dst, c->stream));
    int st = 0;
    HIPCHK(c, hipMemcpyAsync(F, dF, (size_t)n * (m_max + 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&st, dst, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->release(dT);
    c->release(dF);
    c->release(dst);
    return st ? fail(c, QCG_INVALID_ARG, "boys: invalid argument") : QCG_OK;
}

int qcg_rys_roots(qcg_context* c, int nroots, int n, const double* T, double* roots, double* weights) {
    if (!c || n < 0 || (n > 0 && (!T || !roots || !weights))) return QCG_INVALID_ARG;
    if (nroots < 1 || nroots > QCG_RYS_NMAX) return fail(c, QCG_INVALID_ARG, "RysError::InvalidOrder");
    for (int i = 0; i < n; ++i)
        if (!(T[i] >= 0.0)) return fail(c, QCG_INVALID_ARG, "RysError::InvalidArgument");
    if (n == 0) return QCG_OK;
    HIPCHK(c, hipSetDevice(c->device));
    double *dT = nullptr, *dx = nullptr, *dw = nullptr;
    HIPCHK(c, c->alloc(&dT, n));
    HIPCHK(c, c->alloc(&dx, (size_t)n * nroots));
    HIPCHK(c, c->alloc(&dw, (size_t)n * nroots));
    HIPCHK(c, hipMemcpyAsync(dT, T, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_rys_batch(nroots, n, dT, c->d_rys, dx, dw, c->stream));
    HIPCHK(c, hipMemcpyAsync(roots, dx, (size_t)n * nroots * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(weights, dw, (size_t)n * nroots * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->release(dT);
    c->release(dx);
    c->release(dw);
    return QCG_OK;
}

namespace {
struct RhfUser {
    qcg_context* c;
    const double* H;
};
int rhf_fock_cb(void* u, const double* D, double* F, double* e_elec) {
    RhfUser* r = (RhfUser*)u;
    int rc = qcg_build_fock(r->c, r->H, D, F);
    if (rc) return rc;
    const size_t n2 = (size_t)r->c->basis.nbf * r->c->basis.nbf;
    double e = 0.0;  // compute_electronic_energy (scf/mod.rs:898): 1/2 sum D (H + F)
    for (size_t i = 0; i < n2; ++i) e += D[i] * (r->H[i] + F[i]);
    *e_elec = 0.5 * e;
    return QCG_OK;
}
// mixed precision: the last SCF iterations run FP64 builds (scf_driver's refine step)
int refine_fp64(void* u) {
    static_cast<RhfUser*>(u)->c->precision = QCG_PREC_FP64;
    return QCG_OK;
}
int rhf_bound_cb(void* u, const double* D, double* b) { return qcg_error_bound(static_cast<RhfUser*>(u)->c, D, 1.0, b); }
}  // namespace

int qcg_rhf_scf(qcg_context* c, const qcg_scf_config* cfg, const double* D_init, qcg_scf_result* res, double* D,
                double* C, double* eps, double* F) {
    if (!c || !cfg) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    const HostBasis& B = c->basis;
    int nelec = -c->charge;
    for (int z : B.Z) nelec += z;
    if (nelec <= 0 || nelec % 2) return fail(c, QCG_INVALID_ARG, "RHF needs an even electron count");
    const size_t n2 = (size_t)B.nbf * B.nbf;
    std::vector<double> S(n2), T(n2), V(n2), H(n2);
    HIPCHK(c, hipSetDevice(c->device));
    const int r1 = one_electron_device(c, B, S.data(), T.data(), V.data());
    if (r1) return r1;
    for (size_t i = 0; i < n2; ++i) H[i] = T[i] + V[i];
    RhfUser u{c, H.data()};
    qcg_scf_result r{};
    const int prec = c->precision;
    int rc = scf_driver(B.nbf, nelec / 2, S.data(), H.data(), nuclear_repulsion(B), *cfg, rhf_fock_cb, &u, D_init, &r,
                        D, C, eps, F, prec == QCG_PREC_MIXED ? refine_fp64 : nullptr, rhf_bound_cb, &c->scf_history);
    c->precision = prec;
    if (res) *res = r;
    if (rc == QCG_NOT_CONVERGED) return fail(c, rc, "ScfError::NotConverged");
    if (rc == QCG_NUMERICAL) return fail(c, rc, "ScfError::SingularOverlap");
    return rc;
}

}  // extern "C"
