"""Benchmark of the CH-form hot path on one MI355X.

Workload (cfg1): n = 64 qubits, 2000 random Clifford gates from {H, S, CX,
CZ} plus 24 T gates (seeded), compiled to chi stabilizer terms.  One step is
one pass of the hot path over device-resident inputs:
  * the per-term quadratic-form rewrite (k_qform, chi terms), which also
    writes the amplitude kernel's per-term nibble tables,
  * <x|psi> for 65,536 random 64-bit strings (k_amplitude, chi*65536
    CH-form amplitude evaluations),
  * the norm via all chi(chi+1)/2 pairwise CH-form inner products (k_pairs),
the library enqueueing on PyTorch's stream; the host waits for each step's
norm (nx_state_norm, the reference's synchronous call), the amplitudes stay
in device memory.  The same steps fully asynchronous (norm into device
memory, nx_state_norm_async) are timed as well.
value = (amplitude evaluations + inner products) per second of step time.
The timed region records HIP events around the dominant kernel only (its
launch duration gives the roofline figure); a separate loop with events
around every launch gives the per-kernel breakdown.

The same run also measures, by default:
  * cfg1 sampling: sample(range(64), 1024, seed=1001), in the reference's
    FP64 order (the default) and with exact probabilities (exact=True);
  * cfg2 batched vs plain: 256 circuits (n ~ U[16,128], 30 n Cliffords,
    t ~ U[4,32]), norm + 64 Pauli expectations each, as one necstar.Batch
    (autotuned launch shape, captured HIP graph) against the plain per-call
    API (norm() and exp_value(P) for every string of every circuit, all of
    them timed);
  * cfg2 sampling, batched vs plain: the cfg2 circuits with at most 200
    terms, 1024 samples on all qubits each (exact mode), in one
    necstar.sample_many call against one sample() call per circuit;
  * cfg3: the coefficient backward against its forward (bench_cfg3) and
    cfg4: the n=128 amplitude sweep over 2^16 strings (bench_cfg4);
  * the CPU baseline: the oracle (a C restatement of the reference
    algorithm) on a bounded sample of the cfg1 units, on one core and on an
    8-process pool.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "stab-decomp-simulator-rust_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "CH-form amplitude evals/s + inner products/s, n=64; batched vs plain, 1 GPU"
# MI355X: 256 CUs x 64 lanes x 2.4 GHz peak engine clock
CUS, CLOCK = 256, 2.4e9
PEAK_VALU_TOPS = CUS * 64 * CLOCK / 1e12      # int32 lane-ops/s (one VALU wave-instruction per CU per cycle)
PEAK_SALU_GIPS = CUS * CLOCK / 1e9            # scalar instructions/s (one per CU per cycle)
PEAK_LDS_TBS = CUS * 128 * CLOCK / 1e12       # LDS bytes/s (128 B per CU per cycle)


# ---------------------------------------------------------------------------
# CPU baseline (oracle = C restatement of the reference algorithm)

def _cpu_worker(args):
    """Times oracle amplitude evaluations and inner products of the cfg1
    state for about `seconds` (worker `k` of `nw` takes every nw-th unit)."""
    k, nw, seconds = args
    import numpy as np
    import workloads as WL
    from oracle import oracle as O
    cfg = WL.config(1)
    n = cfg["n"]
    t0 = time.perf_counter()
    st = O.State.compile(n, cfg["gates"])
    compile_s = time.perf_counter() - t0
    chi = st.chi
    bits = cfg["bitstrings"]
    t0 = time.perf_counter()
    na, b = 0, k
    while time.perf_counter() - t0 < seconds / 2 and b < len(bits):
        x = np.array([(int(bits[b, 0]) >> q) & 1 for q in range(n)], np.uint8)
        st.amplitude(x)
        na += 1
        b += nw
    t_amp = time.perf_counter() - t0
    t0 = time.perf_counter()
    npairs, i, j = 0, 0, k
    while time.perf_counter() - t0 < seconds / 2:
        st.term(i).inner_product(st.term(j))
        npairs += 1
        j += nw
        while j >= chi:  # next row of the triangle (from the start after the last)
            i, j = (i + 1, i + 1 + (j - chi)) if i < chi - 1 else (0, k)
    t_ip = time.perf_counter() - t0
    return dict(chi=chi, compile_s=compile_s, amp_evals=na * chi, t_amp=t_amp, pairs=npairs, t_ip=t_ip, strings=na)


def cpu_baseline(seconds, workers):
    """One core, then a pool of `workers` processes (an upper bound: the
    reference is single-threaded; rayon is declared but unused)."""
    one = _cpu_worker((0, 1, seconds))
    pool = None
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(workers) as p:
            res = p.map(_cpu_worker, [(k, workers, seconds) for k in range(workers)])
        pool = dict(amp_rate=sum(r["amp_evals"] for r in res) / max(r["t_amp"] for r in res),
                    ip_rate=sum(r["pairs"] for r in res) / max(r["t_ip"] for r in res))
    return one, pool


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


# ---------------------------------------------------------------------------
# committed rocprofv3 counter summaries

def read_counters(kernel):
    """Per-launch figures from profiles/<kernel>_cfg1_counters.txt (HBM
    bytes from FETCH_SIZE / WRITE_SIZE, executed VALU / SALU / LDS
    wave-instructions per unit)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_cfg1_counters.txt")
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        key, _, val = line.partition(":")
        key = key.strip()
        if key in ("hbm_bytes_per_launch", "valu_wave_instr_per_unit", "salu_wave_instr_per_unit",
                   "lds_wave_instr_per_unit", "lds_bytes_per_unit"):
            try:
                out[key] = float(val.split()[0])
            except (ValueError, IndexError):
                pass
    return out


def roofline_for(kernel, units, kms):
    """The kernel's executed work against the issue peaks it is bound by.
    k_pairs: int32 VALU lane-ops (SQ_INSTS_VALU x 64 per pair) against the
    VALU peak, and the scalar-instruction issue (its binding resource);
    k_amplitude: LDS bytes per evaluation (16 nibble-table lookups, 8 B at
    positions 0-7 and 4 B at 8-15, and one 16 B value-table read) against
    the LDS peak."""
    ctr = read_counters(kernel)
    rate = units / (kms / 1e3)
    r = {"kernel": kernel, "units_per_launch": units, "avg_launch_ms": kms, "traffic": ctr.get("hbm_bytes_per_launch")}
    if "valu_wave_instr_per_unit" in ctr:
        valu = ctr["valu_wave_instr_per_unit"] * 64.0
        salu = ctr.get("salu_wave_instr_per_unit", 0.0)
        r["executed"] = {"valu_lane_ops_per_unit": valu, "salu_instr_per_unit": salu,
                         "valu_frac": valu * rate / (PEAK_VALU_TOPS * 1e12),
                         "salu_issue_frac": salu * rate / (PEAK_SALU_GIPS * 1e9)}
    if kernel == "k_amplitude":
        per = ctr.get("lds_bytes_per_unit", 112.0)
        ach = per * rate / 1e12
        r.update({"bound": "lds", "achieved": ach, "peak": PEAK_LDS_TBS, "unit": "TB/s (LDS)",
                  "frac": ach / PEAK_LDS_TBS, "basis": f"{per:g} LDS bytes per (term, string) evaluation"})
    else:
        valu = r.get("executed", {}).get("valu_lane_ops_per_unit")
        if valu is not None:
            ach = valu * rate / 1e12
            r.update({"bound": "valu", "achieved": ach, "peak": PEAK_VALU_TOPS, "unit": "TOP/s (int32 lane-ops)",
                      "frac": ach / PEAK_VALU_TOPS,
                      "basis": "executed VALU lane-ops per pair (rocprofv3 SQ_INSTS_VALU x 64, committed summary); "
                               "the binding resource is scalar issue (executed.salu_issue_frac)"})
    return r


# ---------------------------------------------------------------------------
# cfg2: batched vs plain

def pauli_text(ops):
    return "".join("IXYZ"[x] for x in reversed(ops))


def cfg2_compare(ncirc, steps, log):
    import numpy as np
    import workloads as WL
    from necstar import Batch, PauliString, QuantumState
    from necstar import _lib
    L = _lib.lib()
    cfg = WL.config(2)
    circs = cfg["circuits"][:ncirc]
    t0 = time.perf_counter()
    states = [QuantumState.from_circuit(WL.to_circuit(c["n"], c["gates"])) for c in circs]
    compile_s = time.perf_counter() - t0
    paulis = [[PauliString.from_str(pauli_text(p)) for p in c["paulis"]] for c in circs]
    chis = np.array([s.stabilizer_rank for s in states])
    log(f"cfg2: {len(states)} circuits compiled in {compile_s:.1f} s, chi max {chis.max()}")
    t0 = time.perf_counter()
    b = Batch(states, paulis, autotune=True, graph=True)
    create_s = time.perf_counter() - t0
    J, waves, vw = b.shape
    b.run()  # warm-up replay
    _lib.check(L.nx_profile_enable(1))
    t0 = time.perf_counter()
    for _ in range(steps):
        norms, vals = b.run()
    t1 = time.perf_counter()
    ms, cnt = C.c_double(), C.c_int64()
    _lib.check(L.nx_profile_get(b"batch_pairs", C.byref(ms), C.byref(cnt)))
    _lib.check(L.nx_profile_enable(0))
    batched_s = (t1 - t0) / steps
    log(f"cfg2: batched run {batched_s:.3f} s (J={J}, waves={waves}); plain leg next")
    # the plain per-call path: every call of every circuit, timed
    maxdiff = 0.0
    t0 = time.perf_counter()
    for i, (st, pl) in enumerate(zip(states, paulis)):
        nrm = st.norm()
        ev = np.array([st.exp_value(p) for p in pl])
        maxdiff = max(maxdiff, abs(nrm - norms[i]) / max(abs(nrm), 1e-300))
        nzm = np.abs(ev) > 1e-9
        if np.any(nzm):
            maxdiff = max(maxdiff, float(np.max(np.abs(ev[nzm] - vals[i][nzm]) / np.abs(ev[nzm]))))
        if i % 64 == 63:
            log(f"cfg2: plain leg {i + 1} circuits, {time.perf_counter() - t0:.1f} s")
    plain_s = time.perf_counter() - t0
    nc = len(states)
    pairs = float((chis.astype(float) * (chis + 1) / 2).sum())
    return {
        "workload": f"cfg2: {nc} circuits, n U[16,128], 30n Cliffords, t U[4,32], norm + 64 Pauli expectations "
                    "(weight 2-6) each; sampling not included",
        "circuits": nc, "chi_max": int(chis.max()), "chi_mean": float(chis.mean()),
        "batched_s": batched_s, "batched_circuits_per_s": nc / batched_s,
        "batched_graph_ms": ms.value / max(cnt.value, 1),
        "plain_s": plain_s, "plain_circuits_per_s": nc / plain_s, "plain_calls": nc * 65,
        "batched_vs_plain": plain_s / batched_s, "max_rel_diff_batched_vs_plain": maxdiff,
        "pair_string_values_per_s": pairs * 65 / batched_s,
        "shape": {"J": J, "waves": waves, "variant_words": vw}, "compile_s": compile_s, "batch_create_s": create_s,
    }


def cfg2_sampling(chimax, shots, log, plain=True, chunk=0):
    """cfg2 sampling on the circuits with at most `chimax` terms: `shots`
    Z-basis samples on all qubits per circuit (exact mode), all circuits in
    one necstar.sample_many call against one QuantumState.sample call per
    circuit (both timed in full); the outcomes must be identical.  chunk > 0:
    one sample_many call per `chunk` consecutive circuits, each logged (long
    runs over the large circuits; a state's outcomes do not depend on the
    calls it shares)."""
    import numpy as np
    import necstar
    import workloads as WL
    from necstar import QuantumState
    cfg = WL.config(2)
    sts, qargs, seeds = [], [], []
    for i, c in enumerate(cfg["circuits"]):
        st = QuantumState.from_circuit(WL.to_circuit(c["n"], c["gates"]))
        if st.stabilizer_rank <= chimax:
            sts.append(st)
            qargs.append(list(range(c["n"])))
            seeds.append(20000 + i)
    chis = np.array([s.stabilizer_rank for s in sts])
    log(f"cfg2 sampling: {len(sts)} circuits with chi <= {chimax}, {shots} shots each; batched leg")
    stats = {}
    t0 = time.perf_counter()
    if chunk > 0:
        got = []
        for c0 in range(0, len(sts), chunk):
            part = {}
            got += necstar.sample_many(sts[c0:c0 + chunk], qargs[c0:c0 + chunk], [shots] * len(sts[c0:c0 + chunk]),
                                       seeds[c0:c0 + chunk], stats=part)
            for k, v in part.items():
                stats[k] = stats.get(k, 0) + v
            log(f"cfg2 sampling: circuits {c0}..{min(c0 + chunk, len(sts)) - 1} (chi up to "
                f"{int(chis[c0:c0 + chunk].max())}) done at {time.perf_counter() - t0:.1f} s")
    else:
        got = necstar.sample_many(sts, qargs, [shots] * len(sts), seeds, stats=stats)
    batched_s = time.perf_counter() - t0
    log(f"cfg2 sampling: batched {batched_s:.2f} s ({stats['rounds']} rounds)")
    same, plain_s = None, None
    if plain:
        t0 = time.perf_counter()
        same = True
        for k, st in enumerate(sts):
            same = same and st.sample(qargs[k], shots, seed=seeds[k], exact=True) == got[k]
        plain_s = time.perf_counter() - t0
    return {"workload": f"cfg2 circuits with chi <= {chimax} ({len(sts)} of 256, n 16-128), {shots} Z-basis "
                        "samples on all qubits each, exact mode",
            "circuits": len(sts), "chi_max": int(chis.max()), "chi_sum": int(chis.sum()), "shots": shots,
            "batched_s": batched_s, "plain_s": plain_s,
            "batched_vs_plain": plain_s / batched_s if plain_s is not None else None, "identical_outcomes": same, "rounds": stats["rounds"], "window_passes": stats["windows"],
            "exact_tails": stats["tails"], "round_s": stats["round_s"], "gpu_wait_s": stats["gpu_wait_s"],
            "dfs_wait_s": stats["wait_s"]}


# ---------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default per config: 20, 1, 5, 2)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default per config: 3, 1, 1, 1)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample length per leg (0: skip)")
    ap.add_argument("--cpu-workers", type=int, default=8, help="processes of the pooled CPU baseline (1: skip it)")
    ap.add_argument("--config", type=int, default=1, choices=(1, 2, 3, 4),
                    help="1: amplitude sweep + norm (the metric); 2: batched vs plain alone; "
                         "3: Pauli-loss forward vs backward; 4: n=128 amplitude sweep")
    ap.add_argument("--circuits", type=int, default=256, help="cfg2 circuits (the first N; 0: skip the cfg2 leg)")
    ap.add_argument("--cfg2-steps", type=int, default=2, help="timed batched runs of the cfg2 leg")
    ap.add_argument("--cfg2-sample-chimax", type=int, default=200,
                    help="cfg2 sampling leg: the circuits with at most this many terms (0: skip the leg)")
    ap.add_argument("--cfg2-sample-shots", type=int, default=1024, help="cfg2 sampling leg: shots per circuit")
    ap.add_argument("--cfg2-sample-plain", type=int, default=1, help="cfg2 sampling leg: also time the plain calls")
    ap.add_argument("--cfg2-sample-chunk", type=int, default=0,
                    help="cfg2 sampling leg: circuits per sample_many call, each logged (0: one call)")
    ap.add_argument("--sample-shots", type=int, default=1024,
                    help="also time QuantumState.sample over all 64 qubits with this many shots (cfg1; 0: skip)")
    ap.add_argument("--cfg4-shots", type=int, default=0,
                    help="with --config 4: also sample all 128 qubits with this many shots (exact mode)")
    ap.add_argument("--other-configs", type=int, default=1,
                    help="with --config 1: also run the cfg3 backward/forward and cfg4 amplitude-sweep legs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device copy (PMC calibration)")
    args = ap.parse_args()
    dsteps, dwarm = {1: (20, 3), 2: (1, 1), 3: (5, 1), 4: (2, 1)}[args.config]
    if args.steps is None:
        args.steps = dsteps
    if args.warmup is None:
        args.warmup = dwarm
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(f"# {msg}", file=sys.stderr, flush=True)

    # the CPU baseline first, before this process touches the GPU
    cpu_one = cpu_pool = None
    if args.config == 1 and args.cpu_seconds > 0:
        log("CPU baseline (oracle)")
        cpu_one, cpu_pool = cpu_baseline(args.cpu_seconds, args.cpu_workers)

    import numpy as np
    import torch

    import workloads as WL
    from necstar import QuantumState
    from necstar import _lib

    L = _lib.lib()
    _lib.check(L.nx_init(0))
    if args.config == 3:
        print(json.dumps(bench_cfg3(args.steps, args.warmup)), flush=True)
        return
    if args.config == 2:
        res = cfg2_compare(args.circuits, args.cfg2_steps, log) if args.circuits > 0 else {}
        if args.cfg2_sample_chimax > 0:
            res["sampling"] = cfg2_sampling(args.cfg2_sample_chimax, args.cfg2_sample_shots, log,
                                            plain=args.cfg2_sample_plain, chunk=args.cfg2_sample_chunk)
        print(json.dumps(res), flush=True)
        return
    if args.config == 4:
        print(json.dumps(bench_cfg4(args.steps, args.warmup, args.cfg4_shots, args.copy_probe)), flush=True)
        return
    cfg = WL.config(1)
    n = cfg["n"]
    t0 = time.perf_counter()
    st = QuantumState.from_circuit(WL.to_circuit(n, cfg["gates"]))
    compile_s = time.perf_counter() - t0
    chi = st.stabilizer_rank
    bits = cfg["bitstrings"]
    B = bits.shape[0]
    dev = torch.device("cuda:0")
    d_bits = torch.from_numpy(bits.view(np.int64).copy()).to(dev)
    d_out = torch.empty(2 * B, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    bits_p = C.cast(d_bits.data_ptr(), C.POINTER(C.c_uint64))
    out_p = C.cast(d_out.data_ptr(), C.POINTER(C.c_double))
    ts = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the library enqueues on PyTorch's stream: no cross-stream events per call
    _lib.check(L.nx_set_stream(ts, 1))
    d_norm = torch.zeros(1, dtype=torch.float64, device=dev)
    norm_p = C.cast(d_norm.data_ptr(), C.POINTER(C.c_double))
    norm = C.c_double()

    def step():
        _lib.check(L.nx_state_drop_cache(st._h))
        _lib.check(L.nx_state_amplitudes_async(st._h, B, bits_p, 1, out_p, ts))
        _lib.check(L.nx_state_norm_async(st._h, norm_p, ts))

    def step_sync():
        _lib.check(L.nx_state_drop_cache(st._h))
        _lib.check(L.nx_state_amplitudes_async(st._h, B, bits_p, 1, out_p, ts))
        _lib.check(L.nx_state_norm(st._h, C.byref(norm)))

    def kernel_avg(k):
        ms, cnt = C.c_double(), C.c_int64()
        _lib.check(L.nx_profile_get(k.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value / max(cnt.value, 1), cnt.value

    for _ in range(args.warmup):
        step()
        step_sync()
    torch.cuda.synchronize()
    assert float(d_norm.item()) == norm.value, (float(d_norm.item()), norm.value)
    # per-kernel breakdown: HIP events around every launch of the step (the
    # event records themselves lengthen the step; this loop is not `value`)
    _lib.check(L.nx_profile_only(None))
    _lib.check(L.nx_profile_enable(1))
    torch.cuda.synchronize()
    tb0 = time.perf_counter()
    for _ in range(args.steps):
        step_sync()
    torch.cuda.synchronize()
    ms_step_profiled = (time.perf_counter() - tb0) * 1e3 / args.steps
    prof = {k: kernel_avg(k) for k in ("k_amplitude", "k_pairs", "k_qform", "k_amp_reduce", "k_reduce")}
    _lib.check(L.nx_profile_enable(0))
    dom = "k_pairs" if prof["k_pairs"][0] >= prof["k_amplitude"][0] else "k_amplitude"
    # fully asynchronous steps (norm into device memory), no events
    ts0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_step_async = (time.perf_counter() - ts0) * 1e3 / args.steps
    # the timed region (`value`): synchronous steps, HIP events around the
    # dominant kernel only (its launch duration for the roofline)
    _lib.check(L.nx_profile_only(dom.encode()))
    _lib.check(L.nx_profile_enable(1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step_sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dom_ms, dom_cnt = kernel_avg(dom)
    _lib.check(L.nx_profile_enable(0))
    _lib.check(L.nx_profile_only(None))
    assert dom_cnt == args.steps, (dom, dom_cnt)
    amp = d_out.view(torch.complex128).cpu().numpy()
    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    sampling = None
    if args.sample_shots > 0:
        def kprof(name):
            ms, cnt = C.c_double(), C.c_int64()
            _lib.check(L.nx_profile_get(name.encode(), C.byref(ms), C.byref(cnt)))
            return ms.value, cnt.value
        qs = list(range(n))
        # exact mode: exact node probabilities, certified uniform levels,
        # exact tails (timed, then again with per-launch timing)
        ts0 = time.perf_counter()
        counts = st.sample(qs, args.sample_shots, seed=cfg["sample_seed"], exact=True)
        te = time.perf_counter() - ts0
        _lib.check(L.nx_profile_enable(1))
        p0, s0 = kprof("k_pairs_var"), kprof("k_pairs_sym")
        assert st.sample(qs, args.sample_shots, seed=cfg["sample_seed"], exact=True) == counts
        p1, s1 = kprof("k_pairs_var"), kprof("k_pairs_sym")
        _lib.check(L.nx_profile_enable(0))
        # reference order (the default): every node probability as the
        # reference's FP64 sum, one window pass + fold per 6 levels
        log(f"sampling {args.sample_shots} shots in reference order")
        _lib.check(L.nx_profile_enable(1))
        p2, f2 = kprof("k_pairs_var"), kprof("k_fold")
        ts0 = time.perf_counter()
        counts_ref = st.sample(qs, args.sample_shots, seed=cfg["sample_seed"])
        tr = time.perf_counter() - ts0
        p3, f3 = kprof("k_pairs_var"), kprof("k_fold")
        _lib.check(L.nx_profile_enable(0))
        # DFS nodes with shots (distinct outcome prefixes in the processing
        # order, qubit n-1 first): the reference evaluates 3 norms of the
        # projected state at each (sampling.rs:101-153)
        keys = [k[::-1] for k in counts_ref]
        nodes = sum(len({k[:d] for k in keys}) for d in range(n))
        sampling = {"shots": args.sample_shots, "qubits": n, "dfs_nodes": nodes,
                    "reference_order": {"seconds": tr, "distinct_outcomes": len(counts_ref),
                                        "windows": p3[1] - p2[1], "pass_kernel_ms": p3[0] - p2[0],
                                        "fold_kernel_ms": f3[0] - f2[0]},
                    "exact": {"seconds": te, "distinct_outcomes": len(counts),
                              "uniform_levels": st.uniform_prefix(list(range(n - 1, -1, -1))),
                              "uniform_pass_ms": s1[0] - s0[0], "projection_passes": p1[1] - p0[1],
                              "pass_kernel_ms": p1[0] - p0[0]},
                    "same_outcomes": counts == counts_ref}

    cfg2 = cfg2_compare(args.circuits, args.cfg2_steps, log) if args.circuits > 0 else None
    cfg2_samp = (cfg2_sampling(args.cfg2_sample_chimax, args.cfg2_sample_shots, log)
                 if args.cfg2_sample_chimax > 0 else None)
    cfg3 = cfg4 = None
    if args.other_configs:
        log("cfg3: coefficient backward vs forward")
        cfg3 = bench_cfg3(5, 1)
        log("cfg4: n=128 amplitude sweep")
        cfg4 = bench_cfg4(3, 1)

    ms_step = (t1 - t0) * 1e3 / args.steps
    amp_units = chi * B
    pair_units = chi * (chi + 1) // 2
    value = (amp_units + pair_units) / (ms_step / 1e3)
    amp_ms, pair_ms = prof["k_amplitude"][0], prof["k_pairs"][0]
    amp_rate = amp_units / (amp_ms / 1e3) if amp_ms > 0 else None
    pair_rate = pair_units / (pair_ms / 1e3) if pair_ms > 0 else None
    # the dominant kernel at its duration in the timed region, the other one
    # from the breakdown loop
    roofline = roofline_for(dom, pair_units if dom == "k_pairs" else amp_units, dom_ms)
    other = roofline_for("k_amplitude" if dom == "k_pairs" else "k_pairs",
                         amp_units if dom == "k_pairs" else pair_units, amp_ms if dom == "k_pairs" else pair_ms)

    cpu = None
    if cpu_one is not None:
        t_amp = cpu_one["t_amp"] / max(cpu_one["amp_evals"], 1)
        t_ip = cpu_one["t_ip"] / max(cpu_one["pairs"], 1)
        cpu_step_s = amp_units * t_amp + pair_units * t_ip
        cpu = {"value": (amp_units + pair_units) / cpu_step_s, "unit": "CH-form evals/s", "cores": 1,
               "kind": "port",
               "sample": f"{cpu_one['strings']} strings x {chi} terms amplitudes + {cpu_one['pairs']} pairs of the "
                         f"cfg1 state, oracle (C restatement of the reference algorithm), {args.cpu_seconds:g} s",
               "amplitude_evals_per_s": 1.0 / t_amp, "inner_products_per_s": 1.0 / t_ip,
               "compile_s": cpu_one["compile_s"], "host": host_info()}
        if cpu_pool is not None:
            pool_step = amp_units / cpu_pool["amp_rate"] + pair_units / cpu_pool["ip_rate"]
            cpu["pool"] = {"workers": args.cpu_workers, "value": (amp_units + pair_units) / pool_step,
                           "amplitude_evals_per_s": cpu_pool["amp_rate"], "inner_products_per_s": cpu_pool["ip_rate"],
                           "note": "term/string-parallel upper bound; the reference itself is single-threaded"}
        if sampling is not None:
            # extrapolated, not run: 3 norms of chi(chi+1)/2 pairs per DFS node
            cpu["sampling_estimate_s"] = sampling["dfs_nodes"] * 3 * pair_units * t_ip
            cpu["sampling_estimate"] = "dfs_nodes x 3 norms x chi(chi+1)/2 pairs / oracle inner products per s"

    kernel_sum = sum(v[0] * v[1] for v in prof.values()) / args.steps  # every kernel of the step
    line = {
        "metric": METRIC, "value": value, "unit": "CH-form evals/s (amplitude evals + inner products)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "cfg1: n=64, 2000 Clifford + 24 T gates, 65536 amplitude strings + full norm",
                   "n": n, "chi": chi, "bitstrings": B, "pairs": pair_units, "cliffords": 2000, "t": 24},
        "amplitude_evals_per_s": amp_rate, "inner_products_per_s": pair_rate,
        "kernel_ms": {k: v[0] for k, v in prof.items()}, "kernel_sum_ms": kernel_sum,
        "ms_per_step_profiled": ms_step_profiled,
        "step_overhead_frac": (ms_step_profiled - kernel_sum) / ms_step_profiled,
        "ms_per_step_async": ms_step_async,
        "norm": norm.value, "sum_abs2_amplitudes": float(np.sum(np.abs(amp) ** 2)), "compile_s": compile_s,
        "batched_vs_plain": cfg2["batched_vs_plain"] if cfg2 else None, "cfg2": cfg2,
        "cfg2_sampling": cfg2_samp, "cfg3": cfg3, "cfg4": cfg4,
        "sampling": sampling, "roofline": roofline, "roofline_other": other, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_cfg4(steps, warmup, shots=0, copy_probe=False):
    """Large amplitude sweep: n=128 (two words per row), 6000 Clifford + 40 T
    gates, <x|psi> for 2^16 random 128-bit strings with inputs and outputs
    resident in HBM.  One step is one pass over all chi x 2^16 (term, string)
    evaluations."""
    import numpy as np
    import torch
    import workloads as WL
    from necstar import QuantumState
    from necstar import _lib
    L = _lib.lib()
    cfg = WL.config(4)
    n = cfg["n"]
    t0 = time.perf_counter()
    st = QuantumState.from_circuit(WL.to_circuit(n, cfg["gates"]))
    compile_s = time.perf_counter() - t0
    chi = st.stabilizer_rank
    print(f"# cfg4: chi {chi}, compiled in {compile_s:.1f} s", file=sys.stderr, flush=True)
    bits = cfg["bitstrings"]
    B = bits.shape[0]
    dev = torch.device("cuda:0")
    d_bits = torch.from_numpy(bits.view(np.int64).copy()).to(dev)
    d_out = torch.empty(2 * B, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    bits_p = C.cast(d_bits.data_ptr(), C.POINTER(C.c_uint64))
    out_p = C.cast(d_out.data_ptr(), C.POINTER(C.c_double))
    for _ in range(warmup):
        _lib.check(L.nx_state_amplitudes_device(st._h, B, bits_p, 2, out_p))
    torch.cuda.synchronize()
    _lib.check(L.nx_profile_enable(1))
    t0 = time.perf_counter()
    for _ in range(steps):
        _lib.check(L.nx_state_amplitudes_device(st._h, B, bits_p, 2, out_p))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms, cnt = C.c_double(), C.c_int64()
    _lib.check(L.nx_profile_get(b"k_amplitude", C.byref(ms), C.byref(cnt)))
    _lib.check(L.nx_profile_enable(0))
    step = (t1 - t0) / steps
    evals = float(chi) * B
    kms = ms.value / max(cnt.value, 1)
    amp = d_out.view(torch.complex128).cpu().numpy()
    if copy_probe:  # a 1 GiB device copy: calibrates the HBM byte counters
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
    sampling = "not run (--cfg4-shots 0)"
    if shots > 0:
        # a stated reduced-shot run of the 128-qubit sampling (exact mode: the
        # reference-order records would need 124 GB); per-window progress on stderr
        os.environ["NX_SAMPLE_PROGRESS"] = "1"
        _lib.check(L.nx_profile_enable(1))
        p0 = C.c_double(), C.c_int64()
        _lib.check(L.nx_profile_get(b"k_pairs_var", C.byref(p0[0]), C.byref(p0[1])))
        ts = time.perf_counter()
        counts = st.sample(list(range(n)), shots, seed=1004, exact=True)
        tsamp = time.perf_counter() - ts
        p1 = C.c_double(), C.c_int64()
        _lib.check(L.nx_profile_get(b"k_pairs_var", C.byref(p1[0]), C.byref(p1[1])))
        _lib.check(L.nx_profile_enable(0))
        del os.environ["NX_SAMPLE_PROGRESS"]
        keys = [k[::-1] for k in counts]
        sampling = {"shots": shots, "qubits": n, "mode": "exact", "seconds": tsamp,
                    "distinct_outcomes": len(counts), "window_passes": p1[1].value - p0[1].value,
                    "window_kernel_ms": p1[0].value - p0[0].value,
                    "uniform_levels": st.uniform_prefix(list(range(n - 1, -1, -1))),
                    "dfs_nodes": sum(len({k[:d] for k in keys}) for d in range(n))}
    return {
        "metric": "cfg4 CH-form amplitude evals/s, n=128", "value": evals / step, "unit": "amplitude evals/s",
        "higher_is_better": True, "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": step * 1e3, "k_amplitude_ms": kms, "kernel_evals_per_s": evals / (kms / 1e3),
        "compile_s": compile_s, "sum_abs2": float(np.sum(np.abs(amp) ** 2)),
        "config": {"workload": "cfg4: n=128, 6000 Clifford + 40 T, 2^16 strings of 128 bits", "chi": chi,
                   "bitstrings": B, "sampling": sampling},
    }


def bench_cfg3(steps, warmup):
    """Differentiation workload: n=48, 1500 Clifford + 30 T, 128 Paulis of
    weight 2-6 with N(0,1) weights, coefficients perturbed by N(0, 1e-3).
    Forward: L = sum_k w_k exp_value(P_k) (exp_values: one pass over the
    pairs for all 128 strings); backward: L and dL/dc together
    (grad_exp_values: one pass over tiles of pairs).  On the cfg3 circuit
    every pair value vanishes (each string anticommutes with a stabilizer
    all terms share), so the same is timed on a non-degenerate circuit of the
    same shape (n=48, 30 T, 150 Cliffords, strings with nonzero expectation
    first)."""
    import numpy as np
    import workloads as WL
    from necstar import PauliString, QuantumState

    def ptxt(p):
        return "".join("IXYZ"[x] for x in reversed(p))

    def measure(st, paulis, w):
        for _ in range(warmup):
            st.exp_values(paulis)
            st.grad_exp_values(paulis, w)
        t0 = time.perf_counter()
        for _ in range(steps):
            ev = st.exp_values(paulis)
        t1 = time.perf_counter()
        for _ in range(steps):
            loss, grad = st.grad_exp_values(paulis, w)
        t2 = time.perf_counter()
        fwd, bwd = (t1 - t0) / steps, (t2 - t1) / steps
        chi = st.stabilizer_rank
        return {"chi": chi, "forward_s": fwd, "backward_s": bwd, "backward_over_forward": bwd / fwd,
                "pairs_per_pass": chi * (chi + 1) // 2, "loss": float(np.dot(w, ev)), "loss_bwd": loss,
                "max_abs_grad": float(np.abs(grad).max()),
                "nonzero_expectations": int(np.sum(np.abs(ev) > 1e-12))}

    cfg = WL.config(3)
    n = cfg["n"]
    st = QuantumState.from_circuit(WL.to_circuit(n, cfg["gates"]))
    rng = np.random.default_rng(cfg["perturb_seed"])
    chi = st.stabilizer_rank
    st.set_coefficients(st.coefficients() + 1e-3 * (rng.standard_normal(chi) + 1j * rng.standard_normal(chi)))
    real = measure(st, [PauliString.from_str(ptxt(p)) for p in cfg["paulis"]], cfg["weights"])
    rng = np.random.default_rng(3201)
    st2 = QuantumState.from_circuit(WL.to_circuit(n, WL.random_circuit(rng, n, 150, 30)))
    chi2 = st2.stabilizer_rank
    st2.set_coefficients(st2.coefficients() + 1e-3 * (rng.standard_normal(chi2) + 1j * rng.standard_normal(chi2)))
    cands = WL.random_paulis(rng, n, 600, 1, 3)
    ev = st2.exp_values([PauliString.from_str(ptxt(p)) for p in cands])
    good = [p for p, e in zip(cands, ev) if abs(e) > 1e-3]
    ops = (good + WL.random_paulis(rng, n, 128, 2, 6))[:128]
    nondeg = measure(st2, [PauliString.from_str(ptxt(p)) for p in ops], rng.standard_normal(128))
    return {
        "metric": "coefficient backward / forward cost, cfg3", "value": nondeg["backward_over_forward"],
        "unit": "ratio", "higher_is_better": False, "n_gpus": 1, "steps": steps, "warmup": warmup,
        "nondegenerate": nondeg, "cfg3": real,
        "config": {"workload": "cfg3 shape: n=48, 30 T, 128 Paulis; value on the non-degenerate circuit "
                               "(150 Cliffords), cfg3 itself (1500 Cliffords) beside it"},
    }


if __name__ == "__main__":
    main()
