"""Benchmark: nonlocal stiffness assembly on one MI355X.

Workload (default): unit square h=1/256 (131,072 P1 triangles, 66,049 vertices), delta=0.05,
constant kernel, retriangulate with caps, 7-point outer/inner rules, CG, FP64.  A step is one
complete assembly of the matrix with the mesh already resident in HBM: dual graph, BFS
neighbour lists, integer sparsity pattern, pair quadrature, compaction to CSR.

Prints one JSON line: element-pairs/s (value), ms per assembled matrix, the roofline of the
dominant kernel (algorithmic FP64 flops / its average launch time, measured with HIP events on
the library's stream) and the CPU baseline (the FP64 OpenMP oracle on a bounded sample of the
same workload, on this host's cores).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix are equal on this part)
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (n, delta, kernel dict, config kwargs, jitter)
    "h256_delta0.05_constant": (256, 0.05, {"function": "constant", "horizon": 0.05, "outputdim": 1}, {}, 0.0),
    "h128_delta0.1_parabola_jitter0.2": (128, 0.1, {"function": "parabola", "horizon": 0.1, "outputdim": 1}, {}, 0.2),
    "h64_delta0.2_fractional_s0.4": (64, 0.2, {"function": "fractional", "horizon": 0.2, "outputdim": 1,
                                               "fractional_s": 0.4},
                                     {"close": "fractional", "tensorGaussDegree": 5}, 0.0),
    "h16_delta0.1_constant": (16, 0.1, {"function": "constant", "horizon": 0.1, "outputdim": 1}, {}, 0.0),
}
DEFAULT = "h256_delta0.05_constant"
BATCH = "batch48_microelastic_field"  # 48 meshes of uneven size and horizon in one plan


def algorithmic_flops(st, nPx, nPy):
    """FP64 flop count of the pair quadrature: 150 per outer point (toPhys, retriangulation,
    outer block update) and 40 per inner point (toPhys, pull-back, basis, weight, 4-block
    accumulate).  Interior pairs have one sub-triangle per outer point; cut pairs report the
    sub-triangles they actually integrate."""
    f_int = st["interior"] * (nPx * 150 + nPx * nPy * 40)
    f_cut = st["cut"] * nPx * 150 + st["subtri"] * nPy * 40
    return f_int, f_cut


def cut_kernel_name(conf, kernel):
    """The library's cut-pair kernel with the default plan: k_cut8_o4 for the constant kernel
    with <= 8 outer points and CG (its items integrated from polygon moments), k_cutw (one item
    per lane) for the other kernels and for outdim 2 or DG with <= 16 outer points, else k_cut."""
    nPx = conf["quadrature"]["outer"]["points"].shape[0]
    if (kernel["function"] == "constant" and kernel.get("outputdim", 1) == 1 and nPx <= 8
            and conf.get("ansatz", "CG") == "CG"):
        return "k_cut8_o4"
    return "k_cutw" if nPx <= 16 else "k_cut"


def measured_traffic(kernel_name, workload):
    """HBM bytes per launch of `kernel_name` on `workload`, from the committed counter summary
    profiles/traffic__<kernel>_<workload>_fp64.md (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
    of this bench, each scaled by what the counter reads for a plain copy of known size);
    None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"traffic__{kernel_name}_{workload}_fp64.md")
    try:
        with open(path) as f:
            for line in f:
                m = re.match(r"\s*-?\s*HBM bytes per launch:\s*([0-9.eE+]+)", line)
                if m:
                    return float(m.group(1))
    except OSError:
        pass
    return None


def _cpu_sample(oracle, mesh, kernel, conf, budget_s, nthreads):
    """Oracle run on every stride-th outer element of the whole mesh (interior and boundary
    elements in their true proportion), with the stride calibrated so that the run takes about
    budget_s seconds on nthreads threads (0: all cores)."""
    nE = mesh["nE"]
    stride = max(1, nE // 64)
    while True:
        r = oracle.par_system(mesh, kernel, conf, nthreads=nthreads, stride=stride)
        if r["seconds"] > 1.0 or stride == 1:
            break
        stride = max(1, stride // 8)
    target = max(1, int(stride * r["seconds"] / budget_s))
    if target < stride * 0.7:
        stride = target
        r = oracle.par_system(mesh, kernel, conf, nthreads=nthreads, stride=stride)
    n_outer = (nE + stride - 1) // stride
    return r, stride, n_outer


def cpu_baseline(mesh, kernel, conf, budget_s=15.0, budget_1t_s=8.0):
    """The FP64 OpenMP oracle (the reference's cost structure: per-thread std::map, queue BFS
    with the O(nE) visited reset per outer element, every thread scanning all elements) on a
    strided sample of the whole mesh's outer elements, on all host cores and on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nE = mesh["nE"]
    cores = os.cpu_count() or 1
    r, stride, n_outer = _cpu_sample(oracle, mesh, kernel, conf, budget_s, 0)
    out = {"value": r["n_pairs"] / r["seconds"], "unit": "element-pairs/s", "cores": cores, "kind": "port",
           "cpu_model": _cpu_model(),
           "sample": f"every {stride}-th of {nE} outer elements ({n_outer}), {r['n_pairs']} pairs, "
                     f"{r['seconds']:.2f} s incl. dual graph and merge, OMP threads={cores}"}
    r1, s1, n1 = _cpu_sample(oracle, mesh, kernel, conf, budget_1t_s, 1)
    out["one_thread"] = {"value": r1["n_pairs"] / r1["seconds"], "unit": "element-pairs/s", "cores": 1,
                         "sample": f"every {s1}-th of {nE} outer elements ({n1}), {r1['n_pairs']} pairs, "
                                   f"{r1['seconds']:.2f} s"}
    return out


def _cpu_model():
    """The host CPU model (/proc/cpuinfo), reported with the CPU baseline."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def end_to_end(mesh, kernel, conf, reps=3):
    """Host arrays to host CSR, as a caller of stiffnessMatrix sees it: plan creation (checks,
    estimates, mesh upload H2D, element set-up), the first assembly of the plan (including its
    device allocations), the CSR copy D2H into pageable host memory, and the plan's teardown.
    Medians over reps."""
    from nlfem_amd import nlfem
    rows = []
    for _ in range(reps):
        t0 = time.perf_counter()
        a = nlfem.Assembler(mesh, kernel, conf)
        t1 = time.perf_counter()
        a.assemble()
        t2 = time.perf_counter()
        A = a.csr()
        t3 = time.perf_counter()
        a.close()
        t4 = time.perf_counter()
        rows.append((t1 - t0, t2 - t1, t3 - t2, t4 - t3, t4 - t0))
    med = np.median(np.array(rows) * 1000.0, axis=0)
    h2d = sum(np.asarray(mesh[k]).nbytes for k in ("elements", "elementLabels", "vertices", "vertexLabels"))
    d2h = A.indptr.size * 8 + A.nnz * (4 + 8)
    return {"ms_total": round(float(med[4]), 2), "ms_create_h2d": round(float(med[0]), 2),
            "ms_first_assembly": round(float(med[1]), 2), "ms_csr_d2h": round(float(med[2]), 2),
            "ms_destroy": round(float(med[3]), 2), "h2d_bytes": int(h2d), "d2h_bytes": int(d2h),
            "d2h_GBs": round(d2h / (med[2] / 1000.0) / 1e9, 2), "reps": reps}


def oracle_matrix(mesh, kernel, conf):
    """The FP64 CPU oracle's matrix of the whole mesh (all host cores) and its wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t0 = time.perf_counter()
    r = oracle.par_system(mesh, kernel, conf)
    return r["A"], time.perf_counter() - t0


def error_vs(B, A):
    """Error of the GPU matrix B against the oracle matrix A: relative Frobenius norm, largest
    entry difference relative to max|A|, whether the sparsity patterns are identical, and when
    they are not, how many entries are in one pattern only and the largest of their values
    relative to max|A| (exact zeros are dropped, so an entry that is rounding residue in one
    matrix and exactly zero in the other appears in one pattern only)."""
    import scipy.sparse.linalg as spla
    A = A.tocsr()
    B = B.tocsr()
    A.sort_indices()
    B.sort_indices()
    D = (B - A).tocsr()
    amax = float(abs(A).max()) if A.nnz else 0.0
    dmax = float(abs(D).max()) if D.nnz else 0.0
    out = {"rel_err_fro": float(spla.norm(D) / spla.norm(A)), "max_abs_err_rel": dmax / amax if amax else dmax,
           "pattern_equal": bool(np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices))}
    if not out["pattern_equal"]:
        PA = A.copy(); PA.data[:] = 1.0
        PB = B.copy(); PB.data[:] = 1.0
        X = (PA - PB).tocoo()
        X.eliminate_zeros()
        vals = [abs(A[r, c]) + abs(B[r, c]) for r, c in zip(X.row[:1000], X.col[:1000])]
        out["pattern_diff_entries"] = int(X.nnz)
        out["pattern_diff_max_rel"] = float(max(vals) / amax) if vals and amax else 0.0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS) + [BATCH])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end (host to host) timing")
    ap.add_argument("--no-checks", action="store_true",
                    help="batch: skip the per-mesh error checks against single-mesh assemblies (profiling runs)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the mixed-precision and surrogate+mixed timings of the same workload "
                         "(batch: skip the per-mesh error checks)")
    ap.add_argument("--vs-oracle", action="store_true",
                    help="also assemble the whole workload with the FP64 CPU oracle (all host cores) and report "
                         "each GPU matrix's error against it (batch: the meshes of at most 2,048 elements)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implemented path)", file=sys.stderr)
        sys.exit(2)

    from nlfem_amd import nlfem
    from nlfem_amd.mesh import unit_square
    from nlfem_amd.config import make_configuration

    if args.workload == BATCH:
        return run_batch(args)
    n, delta, kernel, ckw, jitter = WORKLOADS[args.workload]
    mesh = unit_square(n, delta, jitter=jitter, seed=2024)
    conf = make_configuration(**ckw)
    conf["precision"] = args.precision
    print(f"[bench] workload={args.workload} nE={mesh['nE']} building plan", file=sys.stderr, flush=True)
    asm = nlfem.Assembler(mesh, kernel, conf)
    for i in range(args.warmup):
        asm.assemble()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    ms_phase = {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        asm.assemble()
        st = asm.stats()
        for k, v in st.items():
            if k.startswith("ms_"):
                ms_phase[k] = ms_phase.get(k, 0.0) + v
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1000.0 / args.steps
    for k in ms_phase:
        ms_phase[k] /= args.steps
    pairs = st["pairs"]
    value = pairs / (ms_step / 1000.0)
    nPx = conf["quadrature"]["outer"]["points"].shape[0]
    nPy = conf["quadrature"]["inner"]["points"].shape[0]
    f_int, f_cut = algorithmic_flops(st, nPx, nPy)
    dom = "cut" if ms_phase["ms_cut"] >= ms_phase["ms_interior"] else "interior"
    f_dom = f_cut if dom == "cut" else f_int
    t_dom = ms_phase["ms_" + dom] / 1000.0
    achieved = f_dom / t_dom / 1e12
    # the pair quadrature is FP64 vector (VALU) arithmetic, no MFMA: bound by the FP64 vector
    # peak, which on MI355X equals the FP64 matrix peak
    roof = {"bound": "fp64-valu", "compute": "fp64 (VALU; FP64 vector and matrix peaks coincide on MI355X)",
            "kernel": cut_kernel_name(conf, kernel) if dom == "cut" else "k_interior",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": None, "flops_per_launch": f_dom, "ms_per_launch": round(t_dom * 1000, 4)}
    if args.precision == "fp64":
        roof["traffic"] = measured_traffic(roof["kernel"], args.workload)
    out = {
        "metric": "element-pairs/s (assembled ordered element pairs per second, full assembly)",
        "value": value, "unit": "element-pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-sum/f64-accumulate", "data": "synthetic",
        "config": {"workload": args.workload, "nE": mesh["nE"], "nV": mesh["nV"], "delta": delta,
                   "kernel": kernel["function"], "method": "retriangulate+caps", "quadrature": f"{nPx}/{nPy}",
                   "precision": args.precision},
        "pairs": int(pairs), "probes": int(st["probes"]), "nnz": int(st["nnz"]),
        "phases_ms": {k[3:]: round(v, 3) for k, v in ms_phase.items()},
        "roofline": roof,
    }
    A64 = asm.csr()
    Aor = None
    if args.vs_oracle:
        print("[bench] oracle assembly of the whole mesh", file=sys.stderr, flush=True)
        Aor, sec = oracle_matrix(mesh, kernel, conf)
        out["vs_oracle"] = {"oracle_seconds": round(sec, 2), "oracle_threads": os.cpu_count() or 1,
                            args.precision: error_vs(A64, Aor)}
        print(f"[bench] oracle done in {sec:.1f} s", file=sys.stderr, flush=True)
    if not args.no_variants and args.precision == "fp64":
        # the same workload in mixed precision and with the cut-pair surrogate (constant kernel),
        # each compared with the FP64 matrix just assembled
        import scipy.sparse.linalg as spla
        nrm = spla.norm(A64)
        variants = {"mixed": {"precision": "mixed"}}
        if kernel["function"] == "constant":
            variants["surrogate+mixed"] = {"precision": "mixed", "surrogate": {"model": "default"}}
        out["variants"] = {}
        for name, extra in variants.items():
            c2 = dict(conf)
            c2.update(extra)
            print(f"[bench] variant {name}", file=sys.stderr, flush=True)
            a2 = nlfem.Assembler(mesh, kernel, c2)
            for i in range(args.warmup):
                a2.assemble()
            t0 = time.perf_counter()
            for i in range(args.steps):
                a2.assemble()
            ms2 = (time.perf_counter() - t0) * 1000.0 / args.steps
            st2 = a2.stats()
            B = a2.csr()
            a2.close()
            v = {"ms_per_step": ms2, "value": st2["pairs"] / (ms2 / 1000.0), "speedup_vs_fp64": ms_step / ms2,
                 "rel_err_fro_vs_fp64": float(spla.norm(B - A64) / nrm),
                 "pattern_equal": bool(np.array_equal(B.indptr, A64.indptr) and np.array_equal(B.indices, A64.indices)),
                 "ms_cut": round(st2["ms_cut"], 3), "ms_interior": round(st2["ms_interior"], 3)}
            if Aor is not None:
                out["vs_oracle"][name] = error_vs(B, Aor)
            if "surrogate" in extra:
                v["cut_pairs_by_surrogate"] = int(st2["sur_pairs"])
                v["cut_pairs_exact_fallback"] = int(st2["sur_rejected"])
                v["network_items"] = int(st2["sur_items"])
                # BF16 MFMA work of the two networks (12->64->64->64->9: 9,536 multiply-adds per
                # item and network) over the whole cut phase, which also runs the exact kernel on
                # the pairs handed back; the padded per-dispatch count and the counter-based
                # utilisation are in profiles/mfma__k_cut_sur_h256_delta0.05_constant.md
                f = st2["sur_items"] * 2 * 2 * 9536
                v["mfma"] = {"bf16_flops_algorithmic": f, "achieved": round(f / (st2["ms_cut"] / 1000) / 1e12, 2),
                             "peak": 2500.0, "unit": "TFLOP/s",
                             "frac": round(f / (st2["ms_cut"] / 1000) / 2.5e15, 4)}
            out["variants"][name] = v
    asm.close()
    if not args.no_e2e:
        print("[bench] end to end (host arrays to host CSR)", file=sys.stderr, flush=True)
        out["end_to_end"] = end_to_end(mesh, kernel, conf)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(mesh, kernel, conf, args.cpu_budget)
    print(json.dumps(out))


def run_batch(args):
    """48 jittered meshes, h in {1/32..1/256}, delta in [0.05, min(0.2, 12h)], microelastic field
    kernel (outdim 2), 16/7 rules, assembled together: each phase is one launch over all meshes.
    Per-mesh error against the single-mesh FP64 assembly of the same mesh."""
    import scipy.sparse.linalg as spla
    from nlfem_amd import batch, nlfem
    from nlfem_amd.config import make_configuration
    items = batch.c5_meshes(48, seed=2025)
    meshes = [m for m, d, n in items]
    deltas = [d for m, d, n in items]
    kernel = {"function": "linearPrototypeMicroelasticField", "outputdim": 2, "fractional_s": -0.5}
    conf = make_configuration(close="retriangulate", outer="16point", inner="7point")
    print(f"[bench] batch of {len(meshes)} meshes, {sum(m['nE'] for m in meshes)} elements", file=sys.stderr,
          flush=True)
    b = batch.BatchAssembler(meshes, deltas, kernel, conf)
    cut_items, tune_ms = b.autotune()
    print(f"[bench] autotune: {cut_items} items per chunk, cut ms per candidate {tune_ms}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        b.assemble()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        b.assemble()
    ms_step = (time.perf_counter() - t0) * 1000.0 / args.steps
    st = b.stats()
    blocks = b.matrices()
    b.close()
    sur = None
    if not args.no_variants:
        # cut pairs through the field kernel's surrogate (default model), against the exact blocks
        print("[bench] variant surrogate", file=sys.stderr, flush=True)
        bs = batch.BatchAssembler(meshes, deltas, kernel, conf)
        bs.asm.set_cut_items(cut_items)
        bs.set_surrogate("default")
        for i in range(args.warmup):
            bs.assemble()
        t0 = time.perf_counter()
        for i in range(args.steps):
            bs.assemble()
        ms_sur = (time.perf_counter() - t0) * 1000.0 / args.steps
        st2 = bs.stats()
        sblocks = bs.matrices()
        bs.close()
        serr = [float(spla.norm(S - B) / spla.norm(B)) for S, B in zip(sblocks, blocks)]
        num = np.sqrt(sum(spla.norm(S - B) ** 2 for S, B in zip(sblocks, blocks)))
        den = np.sqrt(sum(spla.norm(B) ** 2 for B in blocks))
        sur = {"ms_per_step": ms_sur, "value": st2["pairs"] / (ms_sur / 1000.0), "speedup_vs_fp64": ms_step / ms_sur,
               "ms_cut": round(st2["ms_cut"], 3), "ms_cut_fp64": round(st["ms_cut"], 3),
               "cut_pairs_by_surrogate": int(st2["sur_pairs"]), "cut_pairs_exact_fallback": int(st2["sur_rejected"]),
               "network_items": int(st2["sur_items"]),
               "rel_err_fro_vs_fp64": {"all": float(num / den), "max_per_mesh": max(serr),
                                       "mean_per_mesh": float(np.mean(serr))}}
        print(f"[bench] surrogate: {sur}", file=sys.stderr, flush=True)
    errs, oerrs = [], []
    for (m, d), B in zip(zip(meshes, deltas), blocks):
        if args.no_variants or args.no_checks:  # timing runs (profiles) skip the per-mesh checks
            break
        k = dict(kernel)
        k["horizon"] = d
        a = nlfem.Assembler(m, k, conf)
        A = a.assemble().csr()
        a.close()
        errs.append(float(spla.norm(A - B) / spla.norm(A)))
        if args.vs_oracle and m["nE"] <= 2048:
            Aor, sec = oracle_matrix(m, k, conf)
            e = error_vs(B, Aor)
            e.update({"nE": int(m["nE"]), "delta": round(float(d), 4), "oracle_seconds": round(sec, 2)})
            oerrs.append(e)
            print(f"[bench] oracle mesh nE={m['nE']} delta={d:.3f}: {e}", file=sys.stderr, flush=True)
    print(f"[bench] per-mesh errors done", file=sys.stderr, flush=True)
    out = {
        "metric": "element-pairs/s (assembled ordered element pairs per second, 48 meshes in one plan)",
        "value": st["pairs"] / (ms_step / 1000.0), "unit": "element-pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": BATCH, "meshes": len(meshes), "nE_total": int(sum(m["nE"] for m in meshes)),
                   "h": sorted({n for _, _, n in items}), "delta_range": [min(deltas), max(deltas)],
                   "kernel": kernel["function"], "outdim": 2, "quadrature": "16/7",
                   "cut_items_per_chunk": cut_items, "autotune_cut_ms": {str(k): round(v, 1) for k, v in tune_ms.items()}},
        "pairs": int(st["pairs"]),
        "phases_ms": {k[3:]: round(v, 3) for k, v in st.items() if k.startswith("ms_")},
    }
    if errs:
        out["per_mesh_rel_err_fro"] = {"max": max(errs), "mean": float(np.mean(errs))}
    if sur is not None:
        out["variants"] = {"surrogate": sur}
    if oerrs:
        out["vs_oracle"] = {"meshes": len(oerrs), "oracle_threads": os.cpu_count() or 1,
                            "max_rel_err_fro": max(e["rel_err_fro"] for e in oerrs),
                            "max_abs_err_rel": max(e["max_abs_err_rel"] for e in oerrs),
                            "patterns_equal": all(e["pattern_equal"] for e in oerrs), "per_mesh": oerrs}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
