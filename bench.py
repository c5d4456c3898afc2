"""Benchmark of the CH-form hot path on one MI355X.

Workload (cfg1): n = 64 qubits, 2000 random Clifford gates from {H, S, CX,
CZ} plus 24 T gates (seeded), compiled to chi stabilizer terms.  One step is
one pass of the hot path over device-resident inputs:
  * the per-term quadratic-form rewrite (k_qform, chi terms),
  * <x|psi> for 65,536 random 64-bit strings (k_amplitude, chi*65536
    CH-form amplitude evaluations),
  * the norm via all chi(chi+1)/2 pairwise CH-form inner products (k_pairs).
value = (amplitude evaluations + inner products) per second of step time.

Prints one JSON line.  The CPU baseline is the oracle (a C restatement of
the reference algorithm, one core) on a bounded sample of the same units.
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
# MI355X: 256 CUs x 64 lanes x 2.4 GHz peak engine clock = int32 lane-ops/s
PEAK_VALU_TOPS = 256 * 64 * 2.4e9 / 1e12


def ops_amp(n):  # int32 lane-ops per amplitude evaluation
    w32 = (n + 31) // 32
    return 12 * ((n + 1) // 2) * w32 + 4 * w32


def ops_ip(n):  # int32 lane-ops per CH-form inner product
    return 6 * n * n * ((n + 31) // 32)


def cpu_baseline(cfg, seconds):
    """Oracle (C restatement of the reference, 1 core) on a bounded sample."""
    import numpy as np
    from oracle import oracle as O
    n = cfg["n"]
    t0 = time.perf_counter()
    st = O.State.compile(n, cfg["gates"])
    compile_s = time.perf_counter() - t0
    chi = st.chi
    bits = cfg["bitstrings"]
    # amplitude evaluations: whole strings (all chi terms each)
    t0 = time.perf_counter()
    na = 0
    while time.perf_counter() - t0 < seconds / 2 and na < len(bits):
        x = np.array([(int(bits[na, 0]) >> q) & 1 for q in range(n)], np.uint8)
        st.amplitude(x)
        na += 1
    t_amp = (time.perf_counter() - t0) / (na * chi)
    # pairwise inner products
    t0 = time.perf_counter()
    npairs, i, j = 0, 0, 0
    while time.perf_counter() - t0 < seconds / 2:
        st.term(i).inner_product(st.term(j))
        npairs += 1
        j += 1
        if j >= chi:
            i, j = i + 1, i + 1
    t_ip = (time.perf_counter() - t0) / npairs
    return dict(chi=chi, compile_s=compile_s, t_amp=t_amp, t_ip=t_ip,
                sample=f"{na} strings x {chi} terms amplitudes + {npairs} pairs, cfg1 state")


def read_counters(kernel):
    """Per-launch figures from the committed rocprofv3 --pmc summary of the
    kernel (HBM bytes, VALU / SALU wave instructions per unit)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_cfg1_traffic.txt")
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        key, _, val = line.partition(":")
        if key in ("hbm_bytes_per_launch", "valu_wave_instr_per_unit", "salu_wave_instr_per_unit"):
            out[key] = float(val)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default per config: 20, 1, 5, 2)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default per config: 3, 1, 1, 1)")
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--config", type=int, default=1, choices=(1, 2, 3, 4),
                    help="1: amplitude sweep + norm (the metric); 2: batched vs plain norm + 64 Pauli "
                         "expectations over 256 circuits; 3: Pauli-loss forward vs backward")
    ap.add_argument("--circuits", type=int, default=0, help="cfg2: use only the first N circuits")
    ap.add_argument("--plain-paulis", type=int, default=2,
                    help="cfg2: Pauli strings per circuit timed on the plain path (then scaled to 64)")
    ap.add_argument("--sample-shots", type=int, default=1024,
                    help="also time QuantumState.sample over all 64 qubits with this many shots (cfg1; 0: skip)")
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

    import numpy as np
    import torch

    import workloads as WL
    from necstar import QuantumState
    from necstar import _lib

    L = _lib.lib()
    _lib.check(L.nx_init(0))
    if args.config == 3:
        return bench_cfg3(args)
    if args.config == 2:
        return bench_cfg2(args)
    if args.config == 4:
        return bench_cfg4(args)
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
    norm = C.c_double()

    def step():
        _lib.check(L.nx_state_drop_cache(st._h))
        _lib.check(L.nx_state_amplitudes_device(st._h, B, bits_p, 1, out_p))
        _lib.check(L.nx_state_norm(st._h, C.byref(norm)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(L.nx_profile_enable(1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = {}
    for k in ("k_amplitude", "k_pairs", "k_qform"):
        ms, cnt = C.c_double(), C.c_int64()
        _lib.check(L.nx_profile_get(k.encode(), C.byref(ms), C.byref(cnt)))
        prof[k] = (ms.value / max(cnt.value, 1), cnt.value)
    _lib.check(L.nx_profile_enable(0))
    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    sampling = None
    if args.sample_shots > 0:
        ts = time.perf_counter()
        counts = st.sample(list(range(n)), args.sample_shots, seed=cfg["sample_seed"])
        te = time.perf_counter() - ts
        # again with per-launch timing (synchronizing) for the pass counts
        _lib.check(L.nx_profile_enable(1))
        counts2 = st.sample(list(range(n)), args.sample_shots, seed=cfg["sample_seed"])
        assert counts2 == counts
        pms, pcnt = C.c_double(), C.c_int64()
        _lib.check(L.nx_profile_get(b"k_pairs_var", C.byref(pms), C.byref(pcnt)))
        sms, scnt = C.c_double(), C.c_int64()
        _lib.check(L.nx_profile_get(b"k_pairs_sym", C.byref(sms), C.byref(scnt)))
        _lib.check(L.nx_profile_enable(0))
        # DFS nodes with shots (distinct outcome prefixes in the processing
        # order, qubit n-1 first): the reference evaluates 3 norms of the
        # projected state at each (sampling.rs:101-153)
        keys = [k[::-1] for k in counts]
        nodes = sum(len({k[:d] for k in keys}) for d in range(n))
        sampling = {"shots": args.sample_shots, "qubits": n, "seconds": te, "distinct_outcomes": len(counts),
                    "uniform_levels": st.uniform_prefix(list(range(n - 1, -1, -1))),
                    "uniform_pass_ms": sms.value, "projection_passes": pcnt.value, "pass_kernel_ms": pms.value,
                    "dfs_nodes": nodes}

    ms_step = (t1 - t0) * 1e3 / args.steps
    amp_units = chi * B
    pair_units = chi * (chi + 1) // 2
    value = (amp_units + pair_units) / (ms_step / 1e3)
    amp_ms, pair_ms = prof["k_amplitude"][0], prof["k_pairs"][0]
    amp_rate = amp_units / (amp_ms / 1e3) if amp_ms > 0 else None
    pair_rate = pair_units / (pair_ms / 1e3) if pair_ms > 0 else None
    if pair_ms >= amp_ms:
        kern, per_unit, units, kms = "k_pairs", ops_ip(n), pair_units, pair_ms
    else:
        kern, per_unit, units, kms = "k_amplitude", ops_amp(n), amp_units, amp_ms
    achieved = per_unit * units / (kms / 1e3) / 1e12
    ctr = read_counters(kern)
    roofline = {"bound": "valu", "kernel": kern, "achieved": achieved, "peak": PEAK_VALU_TOPS,
                "unit": "TOP/s (int32 lane-ops)", "frac": achieved / PEAK_VALU_TOPS,
                "ops_per_unit": per_unit, "units_per_launch": units, "avg_launch_ms": kms,
                "ops_basis": ("6 n^2 ceil(n/32) lane-ops per pair: GF(2) elimination of the composed n x n "
                              "matrices" if kern == "k_pairs" else
                              "12 ceil(n/2) ceil(n/32) + 4 ceil(n/32) lane-ops per evaluation: X applied bit "
                              "by bit"),
                "traffic": ctr.get("hbm_bytes_per_launch")}
    # what the kernel executes (rocprofv3 counters of the same launch): the
    # exponential-sum elimination needs fewer operations than ops_basis
    # counts, so frac can exceed 1; these fractions are of the issue peaks
    if "valu_wave_instr_per_unit" in ctr:
        valu = ctr["valu_wave_instr_per_unit"] * 64.0
        salu = ctr.get("salu_wave_instr_per_unit", 0.0)
        rate = units / (kms / 1e3)
        roofline["executed"] = {
            "valu_lane_ops_per_unit": valu, "salu_instr_per_unit": salu,
            "valu_frac": valu * rate / (PEAK_VALU_TOPS * 1e12),
            "salu_issue_frac": salu * rate / (256 * 2.4e9)}

    cpu = None
    if args.cpu_seconds > 0:
        cb = cpu_baseline(cfg, args.cpu_seconds)
        cpu_step_s = amp_units * cb["t_amp"] + pair_units * cb["t_ip"]
        cpu = {"value": (amp_units + pair_units) / cpu_step_s, "unit": "CH-form evals/s", "cores": 1,
               "kind": "port", "sample": cb["sample"],
               "amplitude_evals_per_s": 1.0 / cb["t_amp"], "inner_products_per_s": 1.0 / cb["t_ip"],
               "compile_s": cb["compile_s"]}
        if sampling is not None:
            # extrapolated, not run: 3 norms of chi(chi+1)/2 pairs per DFS node
            # at the measured oracle pair rate
            cpu["sampling_estimate_s"] = sampling["dfs_nodes"] * 3 * pair_units * cb["t_ip"]
            cpu["sampling_estimate"] = "dfs_nodes x 3 norms x chi(chi+1)/2 pairs / oracle inner products per s"

    line = {
        "metric": METRIC, "value": value, "unit": "CH-form evals/s (amplitude evals + inner products)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "cfg1: n=64, 2000 Clifford + 24 T gates, 65536 amplitude strings + full norm",
                   "n": n, "chi": chi, "bitstrings": B, "pairs": pair_units, "cliffords": 2000, "t": 24},
        "amplitude_evals_per_s": amp_rate, "inner_products_per_s": pair_rate,
        "kernel_ms": {k: v[0] for k, v in prof.items()},
        "norm": norm.value, "compile_s": compile_s,
        "batched_vs_plain": None, "sampling": sampling,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def pauli_text(ops):
    return "".join("IXYZ"[x] for x in reversed(ops))


def bench_cfg2(args):
    """Batching workload: 256 circuits, n ~ U[16,128], 30 n Cliffords,
    t ~ U[4,32], each with its norm and 64 Pauli expectations (weight 2-6).
    Batched: one necstar.Batch (autotuned, HIP graph) evaluates all of them
    per run.  Plain: the reference's per-call API, QuantumState.norm() and
    exp_value(P) one string at a time, on every circuit for the first
    --plain-paulis strings; its per-circuit time is norm + 64 x (mean string
    time).  Both legs run in the same process on the same device-resident
    states."""
    import ctypes as C
    import numpy as np
    import workloads as WL
    from necstar import Batch, PauliString, QuantumState
    from necstar import _lib
    L = _lib.lib()
    cfg = WL.config(2)
    circs = cfg["circuits"][: args.circuits] if args.circuits else cfg["circuits"]
    t0 = time.perf_counter()
    states = [QuantumState.from_circuit(WL.to_circuit(c["n"], c["gates"])) for c in circs]
    compile_s = time.perf_counter() - t0
    paulis = [[PauliString.from_str(pauli_text(p)) for p in c["paulis"]] for c in circs]
    chis = np.array([s.stabilizer_rank for s in states])
    ns = np.array([c["n"] for c in circs])
    print(f"# cfg2: {len(states)} circuits compiled in {compile_s:.1f} s, chi max {chis.max()}, "
          f"sum chi^2/2 {float((chis.astype(float) ** 2).sum() / 2):.3e}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = Batch(states, paulis, autotune=True, graph=True)
    create_s = time.perf_counter() - t0
    J, waves, vw = b.shape
    for _ in range(args.warmup):
        b.run()
    _lib.check(L.nx_profile_enable(1))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        norms, vals = b.run()
    t1 = time.perf_counter()
    ms, cnt = C.c_double(), C.c_int64()
    _lib.check(L.nx_profile_get(b"batch_pairs", C.byref(ms), C.byref(cnt)))
    _lib.check(L.nx_profile_enable(0))
    batched_s = (t1 - t0) / args.steps
    graph_ms = ms.value / max(cnt.value, 1)
    print(f"# batched run {batched_s:.3f} s (J={J}, waves={waves})", file=sys.stderr, flush=True)
    # plain per-call path
    K = max(1, min(args.plain_paulis, 64))
    plain_s, maxdiff = 0.0, 0.0
    t_start = time.perf_counter()
    for i, (st, pl) in enumerate(zip(states, paulis)):
        ta = time.perf_counter()
        nrm = st.norm()
        tb = time.perf_counter()
        ev = [st.exp_value(p) for p in pl[:K]]
        tc = time.perf_counter()
        plain_s += (tb - ta) + (tc - tb) / K * len(pl)
        maxdiff = max(maxdiff, abs(nrm - norms[i]) / max(1.0, abs(nrm)),
                      float(np.max(np.abs(np.array(ev) - vals[i][:K]) / np.maximum(1.0, np.abs(ev)))))
    plain_wall = time.perf_counter() - t_start
    nc = len(states)
    pairs = float((chis.astype(float) * (chis + 1) / 2).sum())
    print(json.dumps({
        "metric": "cfg2 batched vs plain, circuits/s (norm + 64 Pauli expectations per circuit)",
        "value": nc / batched_s, "unit": "circuits/s", "higher_is_better": True,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "batched_s": batched_s, "batched_graph_ms": graph_ms, "plain_s_est": plain_s,
        "plain_circuits_per_s": nc / plain_s, "batched_vs_plain": plain_s / batched_s,
        "plain_sample": f"norm + {K} of 64 strings per circuit timed ({plain_wall:.1f} s wall), "
                        f"string time scaled to 64",
        "max_rel_diff_batched_vs_plain": maxdiff,
        "pair_evaluations_per_run": pairs, "pair_string_values_per_s": pairs * 65 / batched_s,
        "shape": {"J": J, "waves": waves, "variant_words": vw}, "compile_s": compile_s, "batch_create_s": create_s,
        "config": {"workload": "cfg2: 256 circuits, n U[16,128], 30n Cliffords, t U[4,32], norm + 64 Paulis "
                               "(weight 2-6) each; sampling not included",
                   "circuits": nc, "chi_max": int(chis.max()), "chi_mean": float(chis.mean()),
                   "n_mean": float(ns.mean())},
    }), flush=True)


def bench_cfg4(args):
    """Large amplitude sweep: n=128 (two words per row), 6000 Clifford + 40 T
    gates, <x|psi> for 2^16 random 128-bit strings with inputs and outputs
    resident in HBM.  One step is one pass over all chi x 2^16 (term, string)
    evaluations."""
    import ctypes as C
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
    for _ in range(args.warmup):
        _lib.check(L.nx_state_amplitudes_device(st._h, B, bits_p, 2, out_p))
    torch.cuda.synchronize()
    _lib.check(L.nx_profile_enable(1))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _lib.check(L.nx_state_amplitudes_device(st._h, B, bits_p, 2, out_p))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms, cnt = C.c_double(), C.c_int64()
    _lib.check(L.nx_profile_get(b"k_amplitude", C.byref(ms), C.byref(cnt)))
    _lib.check(L.nx_profile_enable(0))
    step = (t1 - t0) / args.steps
    evals = float(chi) * B
    kms = ms.value / max(cnt.value, 1)
    amp = d_out.view(torch.complex128).cpu().numpy()
    print(json.dumps({
        "metric": "cfg4 CH-form amplitude evals/s, n=128", "value": evals / step, "unit": "amplitude evals/s",
        "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": step * 1e3, "k_amplitude_ms": kms, "kernel_evals_per_s": evals / (kms / 1e3),
        "compile_s": compile_s, "sum_abs2": float(np.sum(np.abs(amp) ** 2)),
        "config": {"workload": "cfg4: n=128, 6000 Clifford + 40 T, 2^16 strings of 128 bits", "chi": chi,
                   "bitstrings": B, "sampling": "not run (each chi^2 norm is ~3e9 pair evaluations)"},
    }), flush=True)


def bench_cfg3(args):
    """Differentiation workload: n=48, 1500 Clifford + 30 T, 128 Paulis of
    weight 2-6 with N(0,1) weights, coefficients perturbed by N(0, 1e-3).
    Forward: L = sum_k w_k exp_value(P_k); backward: dL/dc (Gram matvec)."""
    import numpy as np
    import workloads as WL
    from necstar import PauliString, QuantumState
    cfg = WL.config(3)
    n = cfg["n"]
    st = QuantumState.from_circuit(WL.to_circuit(n, cfg["gates"]))
    chi = st.stabilizer_rank
    rng = np.random.default_rng(cfg["perturb_seed"])
    c = st.coefficients() + 1e-3 * (rng.standard_normal(chi) + 1j * rng.standard_normal(chi))
    st.set_coefficients(c)
    paulis = [PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in cfg["paulis"]]
    w = cfg["weights"]
    for _ in range(args.warmup):
        st.exp_values(paulis[:4])
        st.grad_exp_values(paulis[:4], w[:4])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev = st.exp_values(paulis)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        loss, grad = st.grad_exp_values(paulis, w)
    t2 = time.perf_counter()
    fwd, bwd = (t1 - t0) / args.steps, (t2 - t1) / args.steps
    pairs = len(paulis) * chi * (chi + 1) // 2
    print(json.dumps({
        "metric": "coefficient backward / forward cost, cfg3", "value": bwd / fwd, "unit": "ratio",
        "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "forward_s": fwd, "backward_s": bwd, "pairs_per_pass": pairs,
        "forward_inner_products_per_s": pairs / fwd, "loss": float(np.dot(w, ev)), "loss_bwd": loss,
        "config": {"workload": "cfg3: n=48, 1500 Clifford + 30 T, 128 Paulis (weight 2-6)", "chi": chi},
    }), flush=True)


if __name__ == "__main__":
    main()
