#!/usr/bin/env python3
"""Fock-build throughput on one MI355X: RHF/6-31G* n-decane C10H22 (194 Cartesian bf).

One step = one integral-direct Fock build F = H + J - K/2 from the converged SCF density
(Schwarz threshold 1e-12; mixed precision: quartets with Q_ab Q_cd < 1e-6 in FP32, the rest FP64,
FP64 accumulation), with H, D and F resident in HBM.  Reported `value` is mixed-precision Fock
builds/s; the same GPU's all-FP64 rate, the speedup, and |dE| between the two energies at that
density are reported beside it.  The CPU baseline is the oracle's restatement of qc-core's
build_fock (scf/mod.rs:825) over the compressed ERI tensor of the same molecule.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK = {"f64": 78.6, "f32": 157.3}  # MI355X dense vector peaks, TFLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ncart(l):
    return (l + 1) * (l + 2) // 2


def kernel_name(la, lb, lc, ld, tile, f64, fused=False):
    """Name of the J/K kernel a class runs under a tile (block 0 = wave-cooperative kernel); fused:
    the mixed-precision launch that runs both parts (jk_kernel_mp / jk_wave_kernel_mp)."""
    nq = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    k = "jk_wave_kernel" if (tile[0] == 0 and nq >= 27) else "jk_kernel"
    if fused:
        return f"{k}_mp<{la}, {lb}, {lc}, {ld}>"
    return f"{k}<{la}, {lb}, {lc}, {ld}, {'double' if f64 else 'float'}>"


def hbm_traffic(name, tile):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc measurement
    (FETCH_SIZE and WRITE_SIZE, scaled by the same counters on a 1 GiB copy), taken with the
    committed tiles; None if absent or if this run's tile for the kernel is another one."""
    f = ROOT / "profiles" / "jk_kernels_decane_631gs_hbm_traffic.json"
    if not f.exists():
        return None
    k = json.loads(f.read_text()).get("kernels", {}).get(name)
    if k is None or list(k.get("tile", [])) != list(tile):
        return None
    return k["bytes_per_launch"]


def grad_model_flops(stats):
    """Algorithmic flops of one ERI-gradient evaluation from the J/K class counts (same Schwarz and
    primitive screens): per surviving primitive quartet n_g [80 + 15 (la+lb+2)(lc+ld+2)
    + 13 N_q (4 + 6 (lb+ld+1))] with n_g = (L+2)//2 + 1 roots, plus 144 N_q per contracted quartet."""
    tot = 0.0
    for x in stats:
        la, lb, lc, ld = x["la"], x["lb"], x["lc"], x["ld"]
        nq = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
        ng = (la + lb + lc + ld + 2) // 2 + 1
        per_prim = ng * (80 + 15 * (la + lb + 2) * (lc + ld + 2) + 13 * nq * (4 + 6 * (lb + ld + 1)))
        tot += per_prim * (x["prim_quartets_fp64"] + x["prim_quartets_fp32"])
        tot += 144 * nq * (x["quartets_fp64"] + x["quartets_fp32"])
    return tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 timed builds (about 0.16 s): run-to-run spread with 20 was about +-3 %
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--fp32-cutoff", type=float, default=1e-6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-gradient", action="store_true", help="skip the gradient evals/s leg")
    ap.add_argument("--no-graph", action="store_true", help="launch every Fock build eagerly")
    ap.add_argument("--no-b3lyp", action="store_true", help="skip the (H2O)16 B3LYP J+K+XC leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-alkane batched SCF leg")
    ap.add_argument("--grad-steps", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tiles instead of autotuning (profiling runs)")
    ap.add_argument("--tiles", default=None,
                    help="JSON {'fp64': tiles, 'mixed': tiles} for the decane build (default: the committed "
                         "profiles/decane_631gs_tiles_tuned.json, which the traffic profile was measured with)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the decane tiles in this run (autotune + whole-build refinement) instead")
    ap.add_argument("--save-tiles", default=None, help="write the decane tiles used to this JSON file")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (5 times): the reference kernel whose "
                         "FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic figures")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = 1 << 27  # 2^27 doubles = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the Fock build runs on a single MI355X; sharding the "
              "quartet work items over several GPUs is designed but not implemented.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from iqcp_amd import workloads
    from iqcp_amd.engine import Engine, ScfConfig

    b = workloads.decane_631gs()
    log(f"workload: C10H22 6-31G* nbf={b.nbf} nshell={b.nshell}")
    eng = Engine(b, device=0)
    tiles = {}
    tiles_file = args.tiles
    committed = ROOT / "profiles" / "decane_631gs_tiles_tuned.json"
    if tiles_file is None and not args.autotune and not args.no_autotune and committed.exists():
        tiles_file = str(committed)
    tiles_source = (f"file {Path(tiles_file).name}" if tiles_file else
                    "library defaults" if args.no_autotune else "autotuned in this run")
    tiles_cfg = json.loads(Path(tiles_file).read_text()) if tiles_file else {}
    costs = {}

    def use_tiles(prec):
        """The tiles of a precision mode from the tiles file, with the per-class launch costs they
        were tuned under (the class kernels' order over the streams) when the file holds them."""
        eng.set_tiles(tiles_cfg[prec])
        eng.set_class_cost(tiles_cfg.get(f"cost_{prec}"))

    if tiles_file:
        # the setup SCF's FP64 builds run the same kernels and tiles as the timed FP64 builds, so a
        # counter run over the whole command sees one tile per kernel
        use_tiles("fp64")
    t0 = time.perf_counter()
    scf = eng.rhf_scf(ScfConfig.tight())
    t_scf = time.perf_counter() - t0
    log(f"setup SCF (FP64, GPU Fock builds): E={scf['energy_total']:.10f} iters={scf['iterations']} {t_scf:.1f}s")
    S, T, V = eng.one_electron()
    H = T + V
    D = scf["D"]
    dev = torch.device("cuda:0")
    dH = torch.tensor(H, device=dev)
    dD = torch.tensor(D, device=dev)
    dF = torch.empty_like(dD)
    stream = torch.cuda.Stream(device=dev)

    def timed(prec):
        eng.set_options(prec, args.fp32_cutoff)
        # per-class tile choice for this precision mode
        if tiles_file:
            use_tiles(prec)
            tiles[prec] = eng.get_tiles()
        elif args.no_autotune:
            eng.set_class_cost(None)
            tiles[prec] = eng.get_tiles()
        else:
            eng.autotune(dD.data_ptr(), 2)
            tiles[prec] = eng.autotune_refine(dD.data_ptr(), 30)
        cost, measured = eng.get_class_cost()
        costs[prec] = cost if measured else None
        eng.set_graph(not args.no_graph)  # one HIP-graph replay per Fock build
        with torch.cuda.stream(stream):
            for _ in range(args.warmup):
                eng.build_fock_device(dH.data_ptr(), dD.data_ptr(), dF.data_ptr(), 0.5, stream.cuda_stream)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(args.steps):
                eng.build_fock_device(dH.data_ptr(), dD.data_ptr(), dF.data_ptr(), 0.5, stream.cuda_stream)
            torch.cuda.synchronize()
            el = time.perf_counter() - t
        eng.set_graph(False)
        F = dF.cpu().numpy()
        E = 0.5 * float(np.sum(D * (H + F))) + b.e_nuc
        return el, F, E

    el64, F64, E64 = timed("fp64")
    elmx, Fmx, Emx = timed("mixed")
    if args.save_tiles:
        out = {p: [[list(t) for t in pair] for pair in tiles[p]] for p in tiles}
        out.update({f"cost_{p}": [round(x, 5) for x in costs[p]] for p in costs if costs[p]})
        Path(args.save_tiles).write_text(json.dumps(out))
    log(f"fp64: {1e3 * el64 / args.steps:.3f} ms/build  mixed: {1e3 * elmx / args.steps:.3f} ms/build")
    # the same build through the host-buffer entry point (qcg_build_fock: D in, F out over PCIe),
    # for the PCIe-inclusive rate; not the headline value
    eng.set_options("mixed", args.fp32_cutoff)
    eng.set_tiles(tiles["mixed"])
    eng.set_class_cost(costs["mixed"])
    eng.build_fock(H, D)
    t = time.perf_counter()
    for _ in range(args.steps):
        eng.build_fock(H, D)
    el_host = time.perf_counter() - t

    # Profiled pass: HIP events around every class kernel on the stream it runs on.
    per_kernel = {}
    for prec in ("fp64", "mixed"):
        eng.set_options(prec, args.fp32_cutoff)
        eng.set_tiles(tiles[prec])
        eng.set_class_cost(costs[prec])
        eng.set_profiling(True)
        acc = None
        with torch.cuda.stream(stream):
            for _ in range(args.steps):
                eng.build_fock_device(dH.data_ptr(), dD.data_ptr(), dF.data_ptr(), 0.5, stream.cuda_stream)
                st = eng.class_stats()
                if acc is None:
                    acc = [dict(s) for s in st]
                else:
                    for a_, s_ in zip(acc, st):
                        a_["ms_fp64"] += s_["ms_fp64"]
                        a_["ms_fp32"] += s_["ms_fp32"]
        eng.set_profiling(False)
        for a_ in acc:
            a_["ms_fp64"] /= args.steps
            a_["ms_fp32"] /= args.steps
        per_kernel[prec] = acc
    for prec in ("fp64", "mixed"):
        log(f"per-class kernels ({prec}): class  quartets64 quartets32  ms64  ms32  model-TFLOP/s")
        for x in per_kernel[prec]:
            c_ = f"({'spd'[x['la']]}{'spd'[x['lb']]}|{'spd'[x['lc']]}{'spd'[x['ld']]})"
            t = (x["ms_fp64"] + x["ms_fp32"]) * 1e-3
            fl = x["flops_fp64"] + x["flops_fp32"]
            log(f"  {c_} {x['quartets_fp64']:10d} {x['quartets_fp32']:10d} {x['ms_fp64']:7.3f} {x['ms_fp32']:7.3f}"
                f" {fl / t / 1e12 if t > 0 else 0:7.3f}")
    # dominant kernel of the mixed build: a class whose FP32 part has quartets but no time of its
    # own ran both parts in one fused launch (its events bracket that launch)
    best = None
    for s in per_kernel["mixed"]:
        fused = s["ms_fp32"] == 0 and s["quartets_fp32"] > 0
        for part, dt in (("fp64", "f64"), ("fp32", "f32")):
            ms = s[f"ms_{part}"]
            if ms > 0 and (best is None or ms > best[0]):
                fl = s["flops_fp64"] + s["flops_fp32"] if fused else s[f"flops_{part}"]
                best = (ms, s, part, dt, fused, fl)
    ms, s, part, dt, fused, fl = best
    ci = per_kernel["mixed"].index(s)
    tile = tiles["mixed"][ci][0 if part == "fp64" else 1]
    kname = kernel_name(s["la"], s["lb"], s["lc"], s["ld"], tile, part == "fp64", fused)
    achieved = fl / (ms * 1e-3) / 1e12
    # FP64 vector ALU bound (a fused launch too: packed FP32 is off, so an FP32 instruction takes an
    # FP64 issue slot): MI355X's FP64 peak is 78.6 TFLOP/s for vector and matrix alike
    roofline = {"bound": "valu", "kernel": kname, "achieved": round(achieved, 3),
                "peak": PEAK["f64" if fused else dt], "unit": "TFLOP/s",
                "frac": round(achieved / PEAK["f64" if fused else dt], 4),
                "traffic": hbm_traffic(kname, tile), "kernel_ms": round(ms, 4), "kernel_flops_model": fl,
                "tile": list(tile), "tiles_source": tiles_source}
    tot_fl = sum(x["flops_fp64"] + x["flops_fp32"] for x in per_kernel["mixed"])
    tot_ms = sum(x["ms_fp64"] + x["ms_fp32"] for x in per_kernel["mixed"])
    kept, n32 = eng.screen_counts()

    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        eng.set_options("fp64")
        eri = eng.eri_compressed()  # the tensor qc-core's build_fock reads (eri_compressed, eri/mod.rs:992)
        # all the host threads this process may use (16 on the GPU box), and one thread (the
        # reference's published protocol is single-threaded, BENCHMARKS.md:23-36); half the
        # time budget each
        try:
            avail = len(os.sched_getaffinity(0))
        except AttributeError:
            avail = os.cpu_count() or 1
        nthr = max(1, min(16, avail))

        def cpu_rate(threads, seconds):
            n_done, t0_ = 0, time.perf_counter()
            Fc_ = None
            while n_done == 0 or time.perf_counter() - t0_ < seconds:
                Fc_ = oracle.build_fock_omp(H, D, eri, b.nbf, threads)
                n_done += 1
            return n_done, time.perf_counter() - t0_, Fc_

        n_all, el, Fc = cpu_rate(nthr, 0.5 * args.cpu_seconds)
        n_one, el1, _ = cpu_rate(1, 0.5 * args.cpu_seconds)
        del eri
        model = "unknown"
        try:
            for line in Path("/proc/cpuinfo").read_text().splitlines():
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(n_all / el, 4), "unit": "Fock builds/s", "cores": nthr, "kind": "port",
               "value_1_thread": round(n_one / el1, 4), "cpu_model": model,
               "sample": f"{n_all} build_fock call(s) on {nthr} threads and {n_one} on 1 thread "
                         f"(scf/mod.rs:825 restated in oracle/) over the "
                         f"{b.nbf * (b.nbf + 1) // 2 * (b.nbf * (b.nbf + 1) // 2 + 1) // 2}-element compressed "
                         f"ERI tensor of the same decane geometry at the same density, {el:.1f}s + {el1:.1f}s",
               "max_abs_F_vs_gpu_fp64": float(np.abs(Fc - F64).max())}
        log(f"cpu baseline: {cpu['value']} builds/s on {nthr} threads, {cpu['value_1_thread']} on 1 thread ({model})")

    # B3LYP leg: one J+K+XC build of (H2O)16/6-31G* on the literal 75 x 302 unpruned grid at its
    # converged KS density, all-FP64 and mixed (mixed: FP32 quartets below the cutoff and the XC
    # grid products in FP32 on MFMA); J/K tiles tuned per precision as for decane.
    ks = None
    if not args.no_b3lyp:
        bw = workloads.water16_631gs()
        ew = Engine(bw, device=0)
        npts = ew.build_grid(75, "standard", False)
        sw = ew.ks_scf(ScfConfig.tight())
        dDw = torch.tensor(sw["D"], device=dev)
        dJw, dKw, dVw = torch.empty_like(dDw), torch.empty_like(dDw), torch.empty_like(dDw)
        dEw = torch.zeros(1, dtype=torch.float64, device=dev)
        ks = {"workload": "B3LYP/6-31G* (H2O)16, J + K + V_xc build at the converged KS density",
              "nbf": bw.nbf, "grid_points": npts, "scf_energy": sw["energy_total"], "scf_iterations": sw["iterations"]}
        for prec in ("fp64", "mixed"):
            ew.set_options(prec, args.fp32_cutoff)
            if not args.no_autotune:
                ew.autotune(dDw.data_ptr(), 2)
                ew.autotune_refine(dDw.data_ptr(), 10)
            # the XC build alone (HIP events on the stream it runs on)
            with torch.cuda.stream(stream):
                ew.xc_device(dDw.data_ptr(), dVw.data_ptr(), dEw.data_ptr(), stream.cuda_stream)
                nx = max(1, args.steps // 4)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(nx):
                    ew.xc_device(dDw.data_ptr(), dVw.data_ptr(), dEw.data_ptr(), stream.cuda_stream)
                e1.record(stream)
                torch.cuda.synchronize()
            ks[f"{prec}_xc_ms"] = round(e0.elapsed_time(e1) / nx, 3)
            with torch.cuda.stream(stream):
                for _ in range(2):
                    ew.build_jk_device(dDw.data_ptr(), dJw.data_ptr(), dKw.data_ptr(), stream.cuda_stream)
                    ew.xc_device(dDw.data_ptr(), dVw.data_ptr(), dEw.data_ptr(), stream.cuda_stream)
                torch.cuda.synchronize()
                nk = max(1, args.steps // 2)
                t = time.perf_counter()
                for _ in range(nk):
                    ew.build_jk_device(dDw.data_ptr(), dJw.data_ptr(), dKw.data_ptr(), stream.cuda_stream)
                    ew.xc_device(dDw.data_ptr(), dVw.data_ptr(), dEw.data_ptr(), stream.cuda_stream)
                torch.cuda.synchronize()
                el_k = time.perf_counter() - t
            Dn = sw["D"]
            Ek = float(np.sum(Dn * (0.5 * dJw.cpu().numpy() - 0.1 * 0.5 * dKw.cpu().numpy()))) + float(dEw.cpu()[0])
            ks[f"{prec}_ms_per_build"] = round(1e3 * el_k / nk, 3)
            ks[f"{prec}_two_electron_plus_xc_energy"] = Ek
        ks["speedup_mixed_vs_fp64"] = round(ks["fp64_ms_per_build"] / ks["mixed_ms_per_build"], 3)
        # XC roofline: 4 G nbf^2 (the two grid GEMMs) + 16 G nbf (rho, grad rho, aow rows) + 300 G
        # (the functional) flops per build on MFMA (FP32 products in the mixed option, FP64 otherwise);
        # algorithmic bytes: Phi and grad Phi read once (4 G nbf elements) + D, V (2 nbf^2 doubles)
        xc_fl = 4.0 * npts * bw.nbf ** 2 + 16.0 * npts * bw.nbf + 300.0 * npts
        for prec, peak, esz in (("mixed", 157.3, 4), ("fp64", 78.6, 8)):
            t_ = ks[f"{prec}_xc_ms"] * 1e-3
            byts = 4.0 * npts * bw.nbf * esz + 2.0 * bw.nbf ** 2 * 8
            ks[f"{prec}_xc_roofline"] = {"bound": "mfma", "achieved": round(xc_fl / t_ / 1e12, 3), "peak": peak,
                                         "unit": "TFLOP/s", "frac": round(xc_fl / t_ / 1e12 / peak, 4),
                                         "flops_model": xc_fl, "algorithmic_GBps": round(byts / t_ / 1e9, 1)}
        ks["dE_mixed_vs_fp64_Ha"] = ks["mixed_two_electron_plus_xc_energy"] - ks["fp64_two_electron_plus_xc_energy"]
        ew.set_options("fp64")
        del ew
        log(f"b3lyp water16: fp64 {ks['fp64_ms_per_build']} ms, mixed {ks['mixed_ms_per_build']} ms per J+K+XC build")

    # Batch leg: 64 n-alkanes (C4..C12, RHF/6-31G), all molecules iterating SCF together, one
    # HIP-graph replay of the batched J/K build per iteration.
    batch = None
    if not args.no_batch:
        from iqcp_amd.engine import BatchEngine
        bb = workloads.alkane_batch_631g()
        be = BatchEngine(bb, device=0)
        cfg_b = ScfConfig.tight(level_shift=0.1)
        if not args.no_autotune:
            # J/K tiles tuned on the converged packed density of an untimed first SCF; the timed
            # SCF below starts again from the core-H guess
            _, D0, _ = be.rhf_scf(cfg_b)
            dD0 = torch.tensor(be.pack(D0), device=dev)
            be.autotune(dD0.data_ptr(), 2)
            del dD0
        t = time.perf_counter()
        res_b, _, jk_ms = be.rhf_scf(cfg_b)
        el_b = time.perf_counter() - t
        fl_b = sum(x["flops_fp64"] + x["flops_fp32"] for x in be.class_stats())
        iters = [r["iterations"] for r in res_b]
        batch = {"workload": "64 n-alkanes C4H10..C12H26 RHF/6-31G, lockstep batched SCF (level shift 0.1 Ha)",
                 "tiles": "default" if args.no_autotune else "autotuned on an untimed first SCF's density",
                 "molecules": len(bb), "nbf_min": min(b.nbf for b in bb), "nbf_max": max(b.nbf for b in bb),
                 "converged": sum(r["converged"] for r in res_b), "batch_iterations": max(iters),
                 "wall_s": round(el_b, 3), "jk_ms_total": round(jk_ms, 2),
                 "jk_ms_per_iteration": round(jk_ms / max(iters), 3),
                 "molecule_scf_per_s": round(len(bb) / el_b, 3)}
        t_b = jk_ms / max(iters) * 1e-3
        batch["roofline"] = {"bound": "valu", "achieved": round(fl_b / t_b / 1e12, 3), "peak": PEAK["f64"],
                             "unit": "TFLOP/s", "frac": round(fl_b / t_b / 1e12 / PEAK["f64"], 4),
                             "flops_model_per_build": fl_b}
        del be
        log(f"batch: {batch['molecules']} molecules, {batch['batch_iterations']} iterations, {batch['wall_s']} s, "
            f"{batch['jk_ms_per_iteration']} ms per batched J/K build")

    # Gradient evals/s: analytic RHF/6-31+G* gradient of n-octane (seeded 0.05 A distortion) at its
    # converged density and energy-weighted density: ERI adjoint kernels on the GPU, one-electron
    # and nuclear terms on the host.
    grad = None
    if not args.no_gradient:
        bo = workloads.octane_631pgs()
        eo = Engine(bo, device=0)
        so = eo.rhf_scf(ScfConfig.tight())
        eo.rhf_gradient(so["D"], so["C"], so["eps"], bo.n_occ)  # warm-up
        ng = max(1, args.grad_steps)
        t = time.perf_counter()
        for _ in range(ng):
            g = eo.rhf_gradient(so["D"], so["C"], so["eps"], bo.n_occ)
        el_g = time.perf_counter() - t
        t = time.perf_counter()
        for _ in range(ng):
            eo.eri_gradient(so["D"], 1.0)
        el_e = time.perf_counter() - t
        fl_g = grad_model_flops(eo.class_stats())
        t_g = el_e / ng
        grad_roof = {"bound": "valu", "achieved": round(fl_g / t_g / 1e12, 3), "peak": PEAK["f64"],
                     "unit": "TFLOP/s", "frac": round(fl_g / t_g / 1e12 / PEAK["f64"], 4),
                     "flops_model_per_eval": fl_g, "of": "the ERI gradient part"}
        grad = {"workload": "RHF/6-31+G* n-octane C8H18 analytic nuclear gradient at the converged density",
                "nbf": bo.nbf, "nshell": bo.nshell, "evals": ng,
                "gradient_evals_per_s": round(ng / el_g, 3), "ms_per_gradient": round(1e3 * el_g / ng, 3),
                "eri_gradient_ms": round(1e3 * el_e / ng, 3),
                "max_abs_gradient": float(np.abs(g).max()),
                "translational_sum": float(np.abs(g.sum(axis=0)).max()), "roofline": grad_roof}
        log(f"gradient: {grad['ms_per_gradient']} ms/eval (ERI part {grad['eri_gradient_ms']} ms) nbf={bo.nbf}")

    out = {
        "metric": "Fock builds/s, mixed vs FP64 on same GPU at |dE|<=1uHa; gradient evals/s",
        "value": round(args.steps / elmx, 3),
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elmx / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (FP32 quartets below Schwarz product 1e-6, FP64 rest and accumulation)",
        "data": "synthetic geometry (ideal all-trans + seeded 0.02 A jitter), converged RHF density",
        "config": {"workload": "RHF/6-31G* n-decane C10H22, integral-direct J+K Fock build at the converged density",
                   "nbf": b.nbf, "nshell": b.nshell, "schwarz_threshold": 1e-12, "fp32_cutoff": args.fp32_cutoff,
                   "unique_quartets_kept": kept, "quartets_fp32": n32},
        "host_buffer_ms_per_step": round(1e3 * el_host / args.steps, 4),
        "fp64_value": round(args.steps / el64, 3),
        "fp64_ms_per_step": round(1e3 * el64 / args.steps, 4),
        "speedup_mixed_vs_fp64": round(el64 / elmx, 3),
        "dE_mixed_vs_fp64_Ha": float(Emx - E64),
        "scf_energy_fp64": scf["energy_total"],
        "scf_iterations": scf["iterations"],
        "all_kernels_model_tflops": round(tot_fl / (tot_ms * 1e-3) / 1e12, 3) if tot_ms > 0 else None,
        "tiles_source": tiles_source,
        "tiles_mixed": tiles.get("mixed"),
        "tiles_fp64": tiles.get("fp64"),
        "b3lyp": ks,
        "batch": batch,
        "gradient": grad,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
