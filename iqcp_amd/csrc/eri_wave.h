// Wave-cooperative shell-quartet evaluation: one 64-lane wavefront works on ONE contracted quartet
// at a time, with the Rys transfer tables in LDS instead of per-lane registers.
//
// The per-lane kernels (one quartet per lane) hold NQ accumulators plus three transfer tables per
// lane; from about NQ = 81 on that no longer fits the register file and the arrays go to scratch.
// Here the work of a quartet is split three ways over the wave:
//   phase A  lanes over (primitive quartet, Rys node, axis): node/weight, 2-D VRR and the 1-D
//            transfer of one axis, written to an LDS table tbl[(axis*HT + h)*ES + e] whose entry
//            index e = (primitive quartet in chunk)*N + node runs fastest (ES odd, so rows of the
//            table start in different LDS banks);  the quadrature weight and the primitive
//            prefactor are folded into the z rows;
//   phase B  lanes over Cartesian components: I(abcd) += Ix Iy Iz over the entries of the chunk;
//   phase C  (caller) digestion of the contracted quartet.
// The integrals are the same as contract_quartet's (eri_device.h), summed in the same order.
#pragma once
#include "eri_device.h"

namespace qcg {

// Memory ordering of LDS among the lanes of one wavefront (no cross-wave synchronisation).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Node r (runtime) of the N-point Rys rule at T: the same fits and asymptotic form as rys_eval.
template <typename R, int N>
__device__ __forceinline__ void rys_eval_one(R T, const R* __restrict__ tab, int r, R& x, R& w) {
    if (T >= (R)qcg_rys_tmax[N]) {
        const R rs = qrsqrt(T);
        const R it = rs * rs;
        x = (R)qcg_rys_hx[rys_tri(N) + r] * it;
        w = (R)qcg_rys_hw[rys_tri(N) + r] * rs;
        return;
    }
    const int k = (int)(T * (R)(1.0 / QCG_RYS_H));
    const R s = (T - (R)(k * QCG_RYS_H)) * (R)(2.0 / QCG_RYS_H) - R(1);
    const R* cx = tab + qcg_rys_off[N] + k * (2 * N * kRysStride) + r * kRysStride;
    const R* cw = cx + N * kRysStride;
    x = poly13(cx, s);
    w = poly13(cw, s);
}

// LDS geometry of one wave for transfer tables with powers up to (TA, TB, TC, TD), N nodes,
// element type R and a budget of kBytes per wave for the table.
template <int TA, int TB, int TC, int TD, int N, typename R, int kBytes = 16384>
struct WaveTable {
    static constexpr int HT = (TA + 1) * (TB + 1) * (TC + 1) * (TD + 1);
    static constexpr int kElems = kBytes / (int)sizeof(R);
    static constexpr int PQ0 = (kElems / (3 * HT) - 1) / N;
    static constexpr int PQ_CAP = PQ0 > 0 ? PQ0 : 1;       // primitive quartets per chunk
    static constexpr int E = PQ_CAP * N;                    // entries per chunk
    static constexpr int ES = E | 1;                        // entry stride (odd)
    static constexpr int ELEMS = 3 * HT * ES;               // table elements of type R
    static constexpr int DOUBLES = (ELEMS * (int)sizeof(R) + 7) / 8;
};

// Phase A for the chunk of primitive quartets [pq0, pq0 + npc) of (P|Q): every lane of the wave
// takes (primitive quartet, node, axis) units.  Primitive pair i of P and j of Q form primitive
// quartet i * Q.npp + j (the loop order of contract_quartet).  Quartets failing the primitive
// screen k_ab k_cd >= 1e-15 get zero z rows.
template <int TA, int TB, int TC, int TD, int N, int ES, typename R>
__device__ __forceinline__ void wave_tables(const PairRec& P, const PairRec& Q, const PrimPair<R>* __restrict__ pp,
                                            const R* __restrict__ rys_tab, int pq0, int npc, R* tbl, int lane) {
    constexpr int HT = (TA + 1) * (TB + 1) * (TC + 1) * (TD + 1);
    constexpr int NB_ = TA + TB, NK_ = TC + TD;
    const R two_pi_52 = (R)(2.0 * kPi * kPi * 1.7724538509055160273);
    const int units = npc * N * 3;
    for (int u = lane; u < units; u += 64) {
        const int axis = u % 3, t = u / 3, r = t % N, pql = t / N;
        const int pq = pq0 + pql;
        const int i = pq / Q.npp, j = pq - i * Q.npp;
        const PrimPair<R> a = pp[P.pp_off + i];
        const PrimPair<R> c = pp[Q.pp_off + j];
        const R p = a.p, q = c.p;
        const R hp = a.hp, hq = c.hp;
        const R rs = qrsqrt(p + q);
        const R ipq = rs * rs;
        const R PQx = a.px - c.px, PQy = a.py - c.py, PQz = a.pz - c.pz;
        const R T = p * q * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
        R x, w;
        rys_eval_one<R, N>(T, rys_tab, r, x, w);
        const R qf = q * ipq, pf = p * ipq;
        const R b00 = R(0.5) * x * ipq;
        const R b10 = hp * (R(1) - x * qf);
        const R b01 = hq * (R(1) - x * pf);
        R PA, QC, PQ, AB, CD;
        if (axis == 0) {
            PA = a.px - (R)P.ax; QC = c.px - (R)Q.ax; PQ = PQx; AB = (R)P.abx; CD = (R)Q.abx;
        } else if (axis == 1) {
            PA = a.py - (R)P.ay; QC = c.py - (R)Q.ay; PQ = PQy; AB = (R)P.aby; CD = (R)Q.aby;
        } else {
            PA = a.pz - (R)P.az; QC = c.pz - (R)Q.az; PQ = PQz; AB = (R)P.abz; CD = (R)Q.abz;
        }
        R g[(NB_ + 1) * (NK_ + 1)];
        vrr2d<NB_, NK_, R>(g, PA - x * qf * PQ, QC + x * pf * PQ, b00, b10, b01);
        R sc = R(1);
        if (axis == 2)
            sc = (prim_kab(a) * prim_kab(c) < kPrimScreen) ? R(0) : w * (two_pi_52 * rs * a.ckp * c.ckp);
        R* row = tbl + axis * HT * ES + pql * N + r;
        hrr1d_put<TA, TB, TC, TD, R>(g, AB, CD, [&](int h, R v) { row[h * ES] = v * sc; });
    }
}

}  // namespace qcg
