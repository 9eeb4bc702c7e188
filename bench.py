#!/usr/bin/env python3
"""Benchmark: MC moves/s of the batched FP64 per-move diagonalisation path.

Workload (the batched FP64 configuration): cubic2d L=16 (N=256 sites),
1024 independent chains, beta_c = 0.5 * 40**(c/1023) (log-spaced in [0.5, 20]),
U_c = {1,2,4,8}[c mod 4], mu_c = mu_f = U/2, f drawn Bernoulli(0.5) per chain,
moves: single-site flip (move_addremove) and nearest-neighbour exchange at
probability 0.5 / 0.5.  A step is one Metropolis move of every chain (proposal,
H assembly, FP64 Householder tridiagonalisation, tridiagonal eigenvalues, logZ,
accept test); the fused kernel runs 16 steps per launch, one workgroup per
(step, chain) item with each chain's steps in order.  Inputs are resident in HBM.

Printed: one JSON line with the contract fields plus
  roofline      FP64 rate of the move-step kernel: (4/3) N^3 flop per eigensolve
                x eigensolves done / kernel time (HIP events on the library
                stream) against MI355X's 78.6 TFLOP/s FP64 peak;
  plain         the single-chain drop-in path (calc_hamiltonian + calc_ed
                through the C ABI with host buffers, one call per move);
  single_chain_cubic2d_L8
                one chain of 10^4 addremove moves at L=8 (N=64) on the device (dense
                FP64, and with rank-one updates), and the plain drop-in at that size;
  measured_sweeps
                sweeps of 16 moves each followed by the on-device measures
                (energy, spectrum mean, spectrum / occupation / IPR histories) of
                all chains, with the overhead over unmeasured sweeps;
  rank1_update, mixed, mixed_cubic3d_L8
                the same chains with rank-one spectral updates / mixed precision,
                and cubic3d L=8 (N=512) mixed against FP64, each with its dOmega
                against FP64 dense chains on the same seeds;
  multi_lattice_chains_triangular
                128 chains on triangular L = 12..24 in one chain set, one persistent
                launch per step block (fkgpu_chains_create_multi);
  adjoint_triangular
                triangular L = 12..24 adjoint gradients (fkgpu_adjoint), configs/s and
                TFLOP/s at (10/3) N^3 flop per configuration, gradients against central
                differences;
  kpm_cubic2d_L64
                cubic2d L=64 (N=4096) Chebyshev-path chains (stochastic-trace logZ,
                local dlogZ of each proposal), with every 100th proposal's dOmega
                against a dense FP64 eigensolve;
  cpu_baseline  the CPU restatement of the reference path (oracle, C++,
                one chain per std::thread) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per move step of the headline move-step kernel on the default workload
# (cubic2d L=16, 1024 chains), measured with rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00
# (scales from a 256 MiB copy, bench.py --copy-calibration) over whole launches; the traffic is
# about proportional to the (step, chain) items.  Keyed by solver variant.
TRAFFIC_BYTES_PER_STEP = {5: 124.5e6 / 16, 10: 1.198e9 / 160}
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of whole launches, scaled by a plain copy, per "
                  "move step x the launch's steps (profiles/k_sweep_q_cubic2d_L16_B1024_hbm_traffic.txt)")
KERNEL_NAMES = {0: "fkk::k_sweep_q<reg_solver<16,4>>", 1: "fkk::k_sweep_q<reg_solver<16,8>>",
                2: "fkk::k_sweep_q<reg_solver<32,8>>", 3: "fkk::k_sweep_q<glob_solver>",
                4: "fkk::k_sweep_q<blocked_solver<double,512>>", 5: "fkk::k_sweep_q<wrow_solver_t<16>>",
                6: "fkk::k_sweep_q<wave_solver>", 7: "fkk::k_sweep_q<blocked_solver<double,256>>",
                8: "fkk::k_sweep_q<blocked_solver<double,1024>>", 9: "fkk::k_sweep_q<blocked_solver<double,512,16>>",
                10: "fkk::k_sweep_q<wrow_solver_t<8>>"}


def workload(L=16, B=1024, seed=20240601):
    N = L * L
    rng = np.random.default_rng(seed)
    f = (rng.random((B, N)) < 0.5).astype(np.int32)
    c = np.arange(B)
    beta = 0.5 * 40.0 ** (c / (B - 1))
    U = np.array([1.0, 2.0, 4.0, 8.0])[c % 4]
    seeds = (seed + c).astype(np.uint32)
    return N, f, beta, U, seeds


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=160)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--chains", type=int, default=1024)
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--plain-moves", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=8, help="mixed-precision timing steps (0: skip)")
    ap.add_argument("--no-c3", action="store_true", help="skip the cubic3d L=8 mixed-vs-FP64 comparison")
    ap.add_argument("--no-c4", action="store_true", help="skip the triangular adjoint workload")
    ap.add_argument("--no-c1", action="store_true", help="skip the single-chain cubic2d L=8 timing")
    ap.add_argument("--green", action="store_true",
                    help="also time G(w,r) / local DOS over recorded eigenvectors (post-processing, off by default)")
    ap.add_argument("--rank1-steps", type=int, default=64, help="rank-one update chain timing steps (0: skip)")
    ap.add_argument("--measure-sweeps", type=int, default=4,
                    help="measured sweeps of 16 moves with on-device measures (0: skip)")
    ap.add_argument("--c5-steps", type=int, default=16,
                    help="cubic2d L=64 stochastic-trace chain timing steps (0: skip)")
    ap.add_argument("--c5-check-steps", type=int, default=25,
                    help="traced L=64 steps whose every 100th proposal is checked against a dense FP64 eigensolve")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB three times (scale of the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copied 256 MiB x3")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); chains are independent, "
              "so several GPUs would each run their own share of the batch.", file=sys.stderr)
        sys.exit(2)

    import fk_mc_amd as fk

    L, B = args.L, args.chains
    N, f, beta, U, seeds = workload(L, B)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f, beta, U, U / 2, U / 2, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])

    # warmup (untimed), then exactly `steps` move steps between synchronisations, as one
    # persistent launch (a launch ends with a tail in which the last chains finish their steps
    # while other CUs idle: 1.6 % of the time of a 16-step launch).  The calls below are
    # generic in the launch size, so the HIP-event time of every launch is known; each call
    # is synchronous, which adds a host round trip per launch.
    ch.set_steps_per_launch(max(args.steps, 1))
    if args.warmup:
        ch.sweep(args.warmup)
    spl = ch.steps_per_launch
    chunks = [spl] * (args.steps // spl) + ([args.steps % spl] if args.steps % spl else [])
    s0 = ch.state()
    launch_ms = []
    t0 = time.perf_counter()
    for k in chunks:
        ch.sweep(k)  # synchronous: returns after the launch completed
        launch_ms.append(ch.last_ms())
    t1 = time.perf_counter()
    s1 = ch.state()
    kern_ms = float(np.sum(launch_ms))
    wall_ms = (t1 - t0) * 1e3
    moves = B * args.steps
    value = moves / (wall_ms / 1e3)
    solves = int((s1["nsolve"] - s0["nsolve"]).sum())
    acc = int((s1["naccept"] - s0["naccept"]).sum())
    flop = solves * (4.0 / 3.0) * N ** 3
    achieved = flop / (kern_ms / 1e3) / 1e12
    full = [ms for k, ms in zip(chunks, launch_ms) if k == spl]
    # HBM bytes per launch at the default workload: the profiled bytes per move step
    # (profiles/, FETCH_SIZE and WRITE_SIZE scaled by a copy kernel) times the mean steps of
    # the timed launches
    per_step_bytes = TRAFFIC_BYTES_PER_STEP.get(ctx.get_solver()) if (L, B) == (16, 1024) else None
    traffic = per_step_bytes * args.steps / len(chunks) if per_step_bytes else None
    roofline = {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                "traffic": traffic,
                "traffic_source": TRAFFIC_SOURCE if traffic else None,
                "traffic_per_step": per_step_bytes,
                "kernel": KERNEL_NAMES.get(ctx.get_solver(), "k_sweep_q"), "steps_per_launch": spl,
                "launches": len(chunks), "steps_per_timed_launch": chunks,
                "kernel_ms_per_launch": round(float(np.mean(full)) if full else kern_ms / len(chunks), 4),
                "kernel_ms_per_launch_all": round(kern_ms / len(chunks), 4),
                "flop_per_launch": flop / len(chunks), "eigensolves_per_launch": solves / len(chunks),
                "work": "(4/3) N^3 FP64 flop per eigensolve (Householder tridiagonalisation), eigensolves "
                        "counted on the device, divided by the HIP-event time of the launches"}

    # plain single-chain drop-in: one calc_hamiltonian+calc_ed per move, host buffers
    plain = None
    if args.plain_moves > 0:
        rng = np.random.default_rng(5)
        fp = f[:1].copy()
        ctx.eig_logz(fp, beta[0], U[0], U[0] / 2)
        tp0 = time.perf_counter()
        for _ in range(args.plain_moves):
            i = rng.integers(N)
            fp[0, i] ^= 1
            ctx.eig_logz(fp, beta[0], U[0], U[0] / 2)
        tp1 = time.perf_counter()
        plain_rate = args.plain_moves / (tp1 - tp0)
        plain = {"moves_per_s": round(plain_rate, 2), "ms_per_move": round((tp1 - tp0) * 1e3 / args.plain_moves, 4),
                 "batched_over_plain": round(value / plain_rate, 1)}

    c1 = None
    if not args.no_c1:
        c1 = c1_block(fk)

    rank1 = None
    if args.rank1_steps > 0:
        rank1 = rank1_block(fk, ctx, f, beta, U, seeds, args.rank1_steps, plain, L)
    measured = None
    if args.measure_sweeps > 0:
        measured = measured_block(fk, ctx, f, beta, U, seeds, args.measure_sweeps)
    green = None
    if args.green:
        green = green_block(fk, ctx, f, beta, U, seeds)
    mixed = None
    if args.mixed_steps > 0:
        mixed = mixed_block(fk, ctx, f, beta, U, seeds, args.mixed_steps, plain)
    c3 = None
    if args.mixed_steps > 0 and not args.no_c3:
        c3 = c3_block(fk, max(2, args.mixed_steps // 2))

    c4 = None
    if not args.no_c4:
        c4 = c4_block(fk)

    multi = None
    if not args.no_c4:
        multi = multi_lattice_block(fk)

    c5 = None
    if args.c5_steps > 0:
        c5 = c5_block(fk, args.c5_steps, args.c5_check_steps)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(L, args.cpu_seconds, ch, ctx, f, beta, U, seeds)

    out = {
        "metric": "MC moves/sec, cubic2d L=16, 1024 batched chains vs plain; dOmega err vs FP64",
        "value": round(value, 2),
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batched FP64 Metropolis, cubic2d L={L} (N={N}), {B} chains, "
                               "beta log-spaced [0.5,20], U in {1,2,4,8}, mu=U/2, f~Bernoulli(0.5), "
                               "moves addremove/nn-exchange 0.5/0.5, one move per chain per step",
                   "lattice": "cubic2d", "L": L, "N": N, "chains": B, "global_batch": B},
        "accept_rate": round(acc / moves, 4),
        "eigensolves_per_s": round(solves / (wall_ms / 1e3), 1),
        "null_moves": round(1.0 - solves / moves, 4),
        "roofline": roofline,
        "plain": plain,
        "single_chain_cubic2d_L8": c1,
        "rank1_update": rank1,
        "measured_sweeps": measured,
        "green_records": green,
        "mixed": mixed,
        "mixed_cubic3d_L8": c3,
        "adjoint_triangular": c4,
        "multi_lattice_chains_triangular": multi,
        "kpm_cubic2d_L64": c5,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ch.close()
    ctx.close()


EPS = np.finfo(np.float64).eps
DOMEGA_ERR_DEF = "|dOmega - dOmega_FP64| / max(|dOmega_FP64|, eps |Omega_FP64|), eps = 2^-52"


def compare_traces(t64, tmx, beta):
    """Per-move dOmega differences while the two chains agree, and the accept mismatches
    (a chain is compared up to its first differing decision).  The relative error is
    taken against max(|dOmega|, eps |Omega|): dOmega is a difference of two O(N) sums,
    so near-zero differences are measured against the rounding of Omega itself."""
    steps, B = t64["acc"].shape
    diffs, rels, mism, ties = [], [], 0, 0
    for c in range(B):
        lz = t64["logz0"][c]
        for s in range(steps):
            if t64["acc"][s, c] != tmx["acc"][s, c]:
                mism += 1
                if abs(np.log(t64["u"][s, c]) - t64["logw"][s, c]) < 1e-9:
                    ties += 1
                break
            lzn = t64["logz"][s, c]
            if np.isfinite(t64["logw"][s, c]) and lzn != 0.0:
                d = abs(tmx["logw"][s, c] - t64["logw"][s, c]) / beta[c]
                d_om = abs(lzn - lz) / beta[c]
                diffs.append(d)
                rels.append(d / max(d_om, EPS * abs(lzn) / beta[c]))
            if t64["acc"][s, c] and lzn != 0.0:
                lz = lzn
    d = np.array(diffs) if diffs else np.zeros(1)
    r = np.array(rels) if rels else np.zeros(1)
    return {"moves_compared": len(diffs), "dOmega_max_abs_diff_vs_fp64": float(d.max()),
            "dOmega_mean_abs_diff_vs_fp64": float(d.mean()), "dOmega_max_rel_err_vs_fp64": float(r.max()),
            "dOmega_err_definition": DOMEGA_ERR_DEF, "accept_mismatches": mism,
            "accept_mismatches_at_near_ties": ties}


def run_chains(fk, ctx, f, beta, U, seeds, precision, steps, trace=False, warmup=2):
    ch = fk.Chains(ctx, f, beta, U, U / 2, U / 2, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.set_precision(precision)
    if trace:
        lz0 = ch.state()["logZ"].copy()
        tr = ch.sweep(steps, trace=True)
        tr["logz0"] = lz0
        ch.close()
        return tr
    if warmup:
        ch.sweep_graph(warmup)
    ch.sweep_graph(steps)
    ms = ch.last_ms() / steps
    nfb = ch.fallbacks()
    ch.close()
    return ms, nfb


def rank1_block(fk, ctx, f, beta, U, seeds, steps, plain, L):
    """The same 1024 chains with rank-one spectral updates (secular equation per
    proposal, MFMA eigenvector update on acceptance, full re-decomposition every 256
    steps): moves/s, and the proposals' dOmega against FP64 dense chains."""
    B = len(beta)
    ch = fk.Chains(ctx, f, beta, U, U / 2, U / 2, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.set_update(fk.UPDATE_RANK1, 256)
    ch.sweep_graph(4)
    ch.sweep_graph(steps)
    ms = ch.last_ms() / steps
    ch.close()
    out = {"moves_per_s": round(B / ms * 1e3, 2), "ms_per_step": round(ms, 4), "steps": steps,
           "refresh_steps": 256}
    if plain:
        out["over_plain"] = round(out["moves_per_s"] / plain["moves_per_s"], 1)
    # dOmega of traced proposals against FP64 dense chains on the same seeds (compared up to
    # each chain's first differing decision)
    nchk, msteps = 64, 40
    sl = slice(0, nchk)
    t64 = run_chains(fk, ctx, f[sl], beta[sl], U[sl], seeds[sl], fk.PRECISION_FP64, msteps, trace=True)
    sub = fk.Chains(ctx, f[sl], beta[sl], U[sl], U[sl] / 2, U[sl] / 2, seeds[sl])
    sub.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    sub.set_update(fk.UPDATE_RANK1, 256)
    tr = sub.sweep(msteps, trace=True)
    sub.close()
    out.update(compare_traces(t64, tr, beta[sl]))
    return out


def measured_block(fk, ctx, f, beta, U, seeds, sweeps, sweep_len=16):
    """The reference's measured schedule on the device: sweeps of sweep_len moves of every
    chain, each followed by measure_energy, measure_spectrum and the spectrum and occupation
    histories of all 1024 chains (records stay on the device until read).  FP64 dense
    chains, and rank-one chains that also record the IPR from their resident eigenvectors.
    The overhead is against the same chains' unmeasured graph replay."""
    B = len(beta)
    out = {"sweep_len": sweep_len, "sweeps": sweeps}
    for name, rank1, flags in (("fp64", False, fk.MEASURE_SPECTRUM_HISTORY | fk.MEASURE_FOCC_HISTORY),
                               ("rank1_ipr", True, fk.MEASURE_SPECTRUM_HISTORY | fk.MEASURE_FOCC_HISTORY
                                | fk.MEASURE_IPR)):
        res = {}
        for measured in (True, False):
            ch = fk.Chains(ctx, f, beta, U, U / 2, U / 2, seeds)
            ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
            if rank1:
                ch.set_update(fk.UPDATE_RANK1, 256)
            if measured:
                ch.measure_enable(sweeps + 1, flags)
                ch.sweep_measure(1, sweep_len)
                ch.sweep_measure(sweeps, sweep_len)
                res["count"] = ch.energies()[0]
            else:
                ch.sweep_graph(sweep_len)
                ch.sweep_graph(sweeps * sweep_len)
            res[measured] = ch.last_ms()
            ch.close()
        out[name] = {"moves_per_s": round(B * sweeps * sweep_len / res[True] * 1e3, 2),
                     "ms_per_sweep": round(res[True] / sweeps, 4),
                     "measure_overhead": round(res[True] / res[False] - 1.0, 4),
                     "measurements": res["count"]}
    return out


def green_block(fk, ctx, f, beta, U, seeds, chains=256, records=4, npts=240):
    """save_gwr's device part over recorded eigenvectors: 256 chains x 4 measurements of the
    headline lattice (1024 records of N x N), G(0 + 0.05i) and the local DOS on the
    reference's 240-point DOS grid.  Kernel time from HIP events; algorithmic FP64 work
    4 n N^3 (G, real and imaginary parts) and 2 n N^2 npts (local DOS)."""
    B = chains
    ch = fk.Chains(ctx, f[:B], beta[:B], U[:B], U[:B] / 2, U[:B] / 2, seeds[:B])
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.measure_enable(records, fk.MEASURE_SPECTRUM_HISTORY | fk.MEASURE_FOCC_HISTORY | fk.MEASURE_EIGENFUNCTIONS)
    ch.sweep_measure(records, 1)
    N = ch.N
    n = B * records
    grid = -6.0 + 12.0 * np.arange(npts) / npts
    t = {}
    for name, call in (("green", lambda: ch.green(0.0, 0.05)), ("local_dos", lambda: ch.local_dos(grid, 0.05))):
        call()
        best = None
        for _ in range(3):
            call()
            ms = ch.last_ms()
            best = ms if best is None else min(best, ms)
        t[name] = best
    ch.close()
    fg = 4.0 * n * N ** 3
    fl = 2.0 * n * N * N * npts
    return {"records": n, "N": N, "npts": npts,
            "green_ms": round(t["green"], 4), "green_tflops": round(fg / t["green"] * 1e-9, 3),
            "green_frac_fp64": round(fg / t["green"] * 1e-9 / 78.6, 4),
            "local_dos_ms": round(t["local_dos"], 4), "local_dos_tflops": round(fl / t["local_dos"] * 1e-9, 3),
            "local_dos_frac_fp64": round(fl / t["local_dos"] * 1e-9 / 78.6, 4)}


def mixed_block(fk, ctx, f, beta, U, seeds, steps, plain):
    """1024 mixed-precision chains (FP32 MFMA reduction + FP64 refinement) on the same
    workload: moves/s, and dOmega / decisions against FP64 chains on the same seeds."""
    B = len(beta)
    ms, nfb = run_chains(fk, ctx, f, beta, U, seeds, fk.PRECISION_MIXED, steps)
    t64 = run_chains(fk, ctx, f, beta, U, seeds, fk.PRECISION_FP64, steps, trace=True)
    tmx = run_chains(fk, ctx, f, beta, U, seeds, fk.PRECISION_MIXED, steps, trace=True)
    out = {"moves_per_s": round(B / ms * 1e3, 2), "ms_per_step": round(ms, 4), "steps": steps,
           "fp64_fallbacks": nfb}
    if plain:
        out["mixed_over_plain"] = round(out["moves_per_s"] / plain["moves_per_s"], 1)
    out.update(compare_traces(t64, tmx, beta))
    return out


def c3_block(fk, steps):
    """cubic3d L=8 (N=512), 256 chains, U in {2,5} x beta in {2,10}, filling ~ U[0.3,0.7]:
    FP64 (blocked MFMA solver) and mixed-precision chains."""
    L, B = 8, 256
    N = L ** 3
    rng = np.random.default_rng(31)
    f = (rng.random((B, N)) < rng.uniform(0.3, 0.7, (B, 1))).astype(np.int32)
    c = np.arange(B)
    U = np.where((c // 64) % 2 == 0, 2.0, 5.0)
    beta = np.where(c // 128 == 0, 2.0, 10.0)
    seeds = (777 + c).astype(np.uint32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 3, L, 1.0, 0.0)
    ms64, _ = run_chains(fk, ctx, f, beta, U, seeds, fk.PRECISION_FP64, steps)
    msmx, nfb = run_chains(fk, ctx, f, beta, U, seeds, fk.PRECISION_MIXED, steps)
    t64 = run_chains(fk, ctx, f, beta, U, seeds, fk.PRECISION_FP64, steps, trace=True)
    tmx = run_chains(fk, ctx, f, beta, U, seeds, fk.PRECISION_MIXED, steps, trace=True)
    out = {"N": N, "chains": B, "steps": steps, "fp64_moves_per_s": round(B / ms64 * 1e3, 2),
           "mixed_moves_per_s": round(B / msmx * 1e3, 2), "fp64_fallbacks": nfb,
           "fp64_solver_variant": ctx.get_solver()}
    out.update(compare_traces(t64, tmx, beta))
    ctx.close()
    return out


def c1_block(fk, nmoves=10000, nplain=2000):
    """The single-chain configuration: cubic2d L=8 (N=64), U=1, beta=1, mu_c = mu_f = 0.5,
    one chain of move_addremove only, 10^4 moves (1 chain x 10^4 steps in 625 launches of
    16 steps, replayed from a graph), dense FP64 and with rank-one updates, and the plain drop-in (calc_hamiltonian + calc_ed
    through the C ABI, host buffers, one call per move) at the same size."""
    L, N = 8, 64
    f = (np.random.default_rng(8).random((1, N)) < 0.5).astype(np.int32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f, 1.0, 1.0, 0.5, 0.5, np.array([1], np.uint32))
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    ch.sweep_graph(nmoves)  # capture + warm-up
    ch.sweep_graph(nmoves)
    ms = ch.last_ms()
    st = ch.state()
    acc = int(st["naccept"][0])
    ch.close()
    # the same chain with rank-one spectral updates (secular equation per move, FP64
    # re-decomposition every 256 steps): same draws, same decisions
    ch = fk.Chains(ctx, f, 1.0, 1.0, 0.5, 0.5, np.array([1], np.uint32))
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    ch.set_update(fk.UPDATE_RANK1, 256)
    ch.sweep_graph(nmoves)
    ch.sweep_graph(nmoves)
    ms_r1 = ch.last_ms()
    st_r1 = ch.state()
    ch.close()
    fp = f.copy()
    rng = np.random.default_rng(9)
    ctx.eig_logz(fp, 1.0, 1.0, 0.5)
    t0 = time.perf_counter()
    for _ in range(nplain):
        fp[0, rng.integers(N)] ^= 1
        ctx.eig_logz(fp, 1.0, 1.0, 0.5)
    plain = nplain / (time.perf_counter() - t0)
    ctx.close()
    return {"N": N, "moves": nmoves, "chain_moves_per_s": round(nmoves / ms * 1e3, 1),
            "chain_us_per_move": round(ms * 1e3 / nmoves, 2), "accept_rate": round(acc / (2 * nmoves), 4),
            "rank1_chain_moves_per_s": round(nmoves / ms_r1 * 1e3, 1),
            "rank1_chain_us_per_move": round(ms_r1 * 1e3 / nmoves, 2),
            "rank1_same_decisions": bool(int(st_r1["naccept"][0]) == acc and np.array_equal(st_r1["f"], st["f"])),
            "rank1_logZ_abs_diff": float(abs(st_r1["logZ"][0] - st["logZ"][0])),
            "plain_moves_per_s": round(plain, 1)}


def _omega_diff(a, b, beta):
    """Omega(a) - Omega(b) for spectra a, b, summed eigenvalue by eigenvalue with
    -log1p(e^{-beta x})/beta differences formed through expm1 (differencing the totals
    would lose ~1e-13 absolute, 1e-7 after division by 2h)."""
    a, b = np.asarray(a), np.asarray(b)
    d = np.empty_like(a)
    pos = b >= 0
    x = np.exp(-beta * b[pos])
    d[pos] = -np.log1p(x * np.expm1(-beta * (a[pos] - b[pos])) / (1 + x)) / beta
    y = np.exp(beta * b[~pos])
    d[~pos] = (a[~pos] - b[~pos]) - np.log1p(y * np.expm1(beta * (a[~pos] - b[~pos])) / (1 + y)) / beta
    return d.sum(axis=-1)


def c4_block(fk, reps=3, per_size=32, nfd=8):
    """Adjoint workload: triangular t = tp = 1, U = 3, beta = 5, mu_c = 1.5, 32 seeded
    configurations each of L = 12, 16, 20, 24 (N = 144 ... 576), one context (and HIP
    stream) per size.  Each size's batch is one fkgpu_adjoint call (calc_ed with
    eigenvectors, rho = V f(Lambda) V^T on the diagonal and the bonds, the gradients) on
    host buffers; the four sizes are issued from four host threads so their launches
    overlap on the device (a 32-configuration batch fills an eighth of the CUs).  Timed on
    the host clock; (10/3) N^3 flop per configuration.  dOmega/dU and dOmega/dt against
    central differences (h = 1e-6) of FP64 spectra on the first configurations of each size."""
    from concurrent.futures import ThreadPoolExecutor
    beta, U, mu, h = 5.0, 3.0, 1.5, 1e-6
    sizes = (12, 16, 20, 24)
    out = {"sizes": {}, "configs": len(sizes) * per_size}
    ctxs, fs, err = {}, {}, 0.0
    for L in sizes:
        fs[L] = (np.random.default_rng(400 + L).random((per_size, L * L)) < 0.5).astype(np.int32)
        ctxs[L] = fk.Context(0).set_lattice(fk.LATTICE_TRIANGULAR, 2, L, 1.0, 1.0)
        ctxs[L].adjoint(fs[L], beta, U, mu)  # warm-up (workspace allocation)

    def one(L):
        t0 = time.perf_counter()
        for _ in range(reps):
            a = ctxs[L].adjoint(fs[L], beta, U, mu)
        return a, (time.perf_counter() - t0) / reps

    seq = {L: (one(L), one(L))[1] for L in sizes}  # each size alone (the first pass a warm-up)
    with ThreadPoolExecutor(len(sizes)) as ex:
        t0 = time.perf_counter()
        res = dict(zip(sizes, ex.map(one, sizes)))
        wall = (time.perf_counter() - t0) / reps
    flop_total = 0.0
    for L in sizes:
        N = L * L
        a, _ = res[L]
        f = fs[L][:nfd]
        fdU = _omega_diff(ctxs[L].eig_logz(f, beta, U + h, mu)["evals"],
                          ctxs[L].eig_logz(f, beta, U - h, mu)["evals"], beta) / (2 * h)
        spec = {}
        for t in (1.0 + h, 1.0 - h):
            c2 = fk.Context(0).set_lattice(fk.LATTICE_TRIANGULAR, 2, L, t, 1.0)
            spec[t] = c2.eig_logz(f, beta, U, mu)["evals"]
            c2.close()
        fdt = _omega_diff(spec[1.0 + h], spec[1.0 - h], beta) / (2 * h)
        e = max(np.abs(a["dOmega_dU"][:nfd] - fdU).max(), np.abs(a["dOmega_dt"][:nfd] - fdt).max())
        flop = per_size * (10.0 / 3.0) * N ** 3
        sec = seq[L][1]
        out["sizes"][f"L{L}"] = {"N": N, "ms_per_batch_alone": round(sec * 1e3, 3),
                                 "configs_per_s_alone": round(per_size / sec, 1),
                                 "grad_max_abs_err_vs_fd": float(e)}
        flop_total += flop
        err = max(err, e)
        ctxs[L].close()
    out["ms_all_sizes_concurrent"] = round(wall * 1e3, 3)
    out["configs_per_s"] = round(len(sizes) * per_size / wall, 1)
    out["grad_max_abs_err_vs_fd"] = float(err)
    out["roofline"] = {"bound": "mfma", "achieved": round(flop_total / wall / 1e12, 4), "peak": FP64_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": round(flop_total / wall / 1e12 / FP64_PEAK_TFLOPS, 5),
                       "work": "(10/3) N^3 FP64 flop per configuration, host-clock time of the four "
                               "concurrent batched calls"}
    return out


def exact_logz_2d(L, f, beta, U, mu):
    """logZ of cubic2d (t=1, periodic) H = hop + diag(-mu + U f) from a dense FP64
    eigensolve (LAPACK through numpy): the exact value the stochastic trace estimates."""
    n = L * L
    H = np.zeros((n, n))
    idx = np.arange(n).reshape(L, L)
    for shift in ((1, 0), (0, 1)):
        j = np.roll(idx, (-shift[0], -shift[1]), axis=(0, 1)).ravel()
        H[idx.ravel(), j] -= 1.0
        H[j, idx.ravel()] -= 1.0
    H[np.arange(n), np.arange(n)] = -mu + U * f
    ev = np.linalg.eigvalsh(H)
    return float(np.logaddexp(0.0, -beta * ev).sum())


def c5_block(fk, steps, check_steps, every=100):
    """cubic2d L=64 (N=4096), 16 chains, beta=10, U=2, mu=U/2 on the Chebyshev path:
    the chains' logZ from the stochastic trace (R=64 Rademacher probes per chain, M=256
    moments, Jackson kernel, FP32 recursion with FP64 moment sums), each proposal's dlogZ
    from the local estimator (FP64 Chebyshev vectors started at the changed sites and a
    Matsubara sum): moves/s, and dOmega of every 100th proposal against an exact FP64
    dense eigensolve of the old and new configuration."""
    L, B, R, M = 64, 16, 64, 256
    N = L * L
    beta, U = 10.0, 2.0
    rng = np.random.default_rng(64)
    f0 = (rng.random((B, N)) < 0.5).astype(np.int32)
    seeds = (6400 + np.arange(B)).astype(np.uint32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ones = np.ones(B)
    ch = fk.Chains(ctx, f0, beta * ones, U * ones, U / 2 * ones, U / 2 * ones, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.set_kpm(R=R, M=M, seed=17, jackson=True)
    est = ch.kpm_delta()
    out = {"N": N, "chains": B, "R": R, "M": M, "beta": beta, "U": U,
           "dlogZ_estimator": {"kind": "local Chebyshev moments + Matsubara sum (FP64)",
                               "moments": est["moments"], "frequencies": est["frequencies"]}}
    # validation pass (traced): replay each chain's accepted moves to recover the configurations
    if check_steps > 0:
        lz = ch.state()["logZ"].copy()
        tr = ch.sweep(check_steps, trace=True)
        f = f0.copy()
        errs, checks, t_ed = [], [], 0.0
        due = 0
        for s in range(check_steps):
            for c in range(B):
                k = s * B + c
                solved = tr["logz"][s, c] != 0.0
                a, b = tr["a"][s, c], tr["b"][s, c]
                g = f[c].copy()
                if tr["kind"][s, c] == fk.MOVE_HOP and b >= 0:
                    g[a], g[b] = f[c, b], f[c, a]
                elif a >= 0:
                    g[a] ^= 1
                if solved and k >= due:
                    t = time.perf_counter()
                    d_ex = -(exact_logz_2d(L, g, beta, U, U / 2) - exact_logz_2d(L, f[c], beta, U, U / 2)) / beta
                    t_ed += time.perf_counter() - t
                    d_kpm = -(tr["logz"][s, c] - lz[c]) / beta
                    errs.append(abs(d_kpm - d_ex))
                    checks.append({"move": k, "dOmega_kpm": round(d_kpm, 6), "dOmega_exact": round(d_ex, 6)})
                    due = (k // every + 1) * every
                if tr["acc"][s, c]:
                    f[c] = g
                    if solved:
                        lz[c] = tr["logz"][s, c]
        out.update({"checked_moves": len(errs), "dOmega_max_abs_err_vs_dense_fp64": float(max(errs, default=0.0)),
                    "dOmega_mean_abs_err_vs_dense_fp64": float(np.mean(errs)) if errs else 0.0,
                    "checks": checks, "dense_check_seconds": round(t_ed, 2),
                    "accept_rate_traced": round(float(tr["acc"].mean()), 4)})
    # timed pass: graph-captured steps
    ch.sweep_graph(2)
    ns0 = ch.state()["nsolve"].copy()
    ch.sweep_graph(steps)
    ms = ch.last_ms() / steps
    st = ch.state()
    ch.close()
    ctx.close()
    nnz = 4 * N
    solved = float((st["nsolve"] - ns0).sum()) / steps  # proposals weighed per step (null moves skip)
    # each weighed proposal: one or two FP64 Chebyshev vectors of Mloc - 1 stencil steps,
    # 2 (nnz + 2N) flop per step and vector (the two-site share of the move mix is ~2/3)
    vec = 1.0 + 2.0 / 3.0
    flop = solved * vec * (est["moments"] - 1) * 2 * (nnz + 2 * N)
    achieved = flop / (ms / 1e3) / 1e12
    out.update({"steps": steps, "moves_per_s": round(B / ms * 1e3, 2), "ms_per_step": round(ms, 4),
                "proposals_weighed_per_step": round(solved, 2),
                "roofline": {"bound": "lds", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                             "work": "(Mloc - 1) 2 (nnz + 2N) FP64 flop per Chebyshev vector, 1 + 2/3 vectors per "
                                     "weighed proposal, whole step time; one vector's recursion is one workgroup "
                                     "whose steps are bound by LDS reads of the neighbour values"},
                "logZ_finite": bool(np.all(np.isfinite(st["logZ"])))})
    return out


def multi_lattice_block(fk, steps=8):
    """128 chains of uneven sizes in one chain set: triangular (t = tp = 1) L in {12, 16, 20, 24}
    (N = 144, 256, 400, 576), 32 chains each, U = 3, beta = 5, mu = U/2, addremove /
    nearest-neighbour exchange 0.5 / 0.5, every chain's move steps in one persistent launch
    (the blocked FP64 solver serves all sizes); beside it the four sizes as four chain sets
    run one after the other."""
    Ls = [12, 16, 20, 24]
    ctxs = [fk.Context(0).set_lattice(fk.LATTICE_TRIANGULAR, 2, L, 1.0, 1.0) for L in Ls]
    Ns = [c.N for c in ctxs]
    nmax = max(Ns)
    per = 32
    B = per * len(Ls)
    lat = np.repeat(np.arange(len(Ls)), per)
    rng = np.random.default_rng(24)
    f = np.zeros((B, nmax), np.int32)
    for c in range(B):
        f[c, :Ns[lat[c]]] = (rng.random(Ns[lat[c]]) < 0.5)
    beta, U = 5.0 * np.ones(B), 3.0 * np.ones(B)
    seeds = (2400 + np.arange(B)).astype(np.uint32)
    moves = [(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)]
    ch = fk.Chains(ctxs, f, beta, U, U / 2, U / 2, seeds, lattice=lat)
    ch.set_moves(moves)
    ch.sweep(2)
    s0 = ch.state()
    ch.sweep(steps)
    ms = ch.last_ms()
    s1 = ch.state()
    ch.close()
    solves = (s1["nsolve"] - s0["nsolve"]).astype(float)
    flop = float(np.sum(solves * (4.0 / 3.0) * np.array(Ns)[lat] ** 3))
    sep_ms = 0.0
    for li, L in enumerate(Ls):
        idx = lat == li
        one = fk.Chains(ctxs[li], f[idx, :Ns[li]], beta[idx], U[idx], U[idx] / 2, U[idx] / 2, seeds[idx])
        one.set_moves(moves)
        one.sweep(2)
        one.sweep(steps)
        sep_ms += one.last_ms()
        one.close()
    for c in ctxs:
        c.close()
    achieved = flop / (ms / 1e3) / 1e12
    return {"lattices": "triangular L=12,16,20,24 (N=144,256,400,576), 32 chains each", "chains": B,
            "steps": steps, "one_launch_ms_per_step": round(ms / steps, 3),
            "moves_per_s": round(B * steps / (ms / 1e3), 1),
            "four_chain_sets_in_turn_ms_per_step": round(sep_ms / steps, 3),
            "roofline": {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                         "work": "(4/3) N_c^3 FP64 flop per eigensolve of chain c, device solve counts, HIP-event time"}}


def cpu_baseline(L, seconds, ch, ctx, f, beta, U, seeds):
    """Oracle (the C++ restatement of the reference CPU path: Householder + implicit QR per
    proposal, FP64, libstdc++ draws) on the host cores, on the batched FP64 workload: one
    chain per thread, the chains spread evenly over the batch's beta / U mix, addremove /
    nearest-neighbour exchange 0.5 / 0.5; a bounded sample of about `seconds`.  Beside it
    the single-chain cubic2d L=8 case and a cubic3d L=8 sample, and the dOmega of a short
    traced device sweep against the oracle chain of the same seeds."""
    from oracle import oracle as orc
    nproc = os.cpu_count() or 1
    # one GPU's share of the box's CPUs (16 threads per GPU) bounds the worker pool
    threads = max(1, min(nproc, 16))
    B = len(beta)
    idx = np.round(np.linspace(0, B - 1, threads)).astype(int)
    moves = [(orc.MOVE_ADDREMOVE, 0.5), (orc.MOVE_HOP, 0.5)]

    def run(sel, nmoves):
        t = time.perf_counter()
        tot, ns = orc.bench_chain_set(orc.KIND_HYPERCUBIC, 2, L, beta[sel], U[sel], U[sel] / 2, f[sel], seeds[sel],
                                      moves, nmoves, t=1.0, tp=0.0)
        return tot, ns, time.perf_counter() - t

    # single core first (also calibrates the sample size)
    tot1, ns1, dt1 = run(idx[:1], 24)
    tot1, ns1, dt1 = run(idx[:1], max(24, int(min(3.0, seconds / 4) / (dt1 / 24))))
    per_move = dt1 / tot1
    nmoves = max(8, int(seconds / per_move))
    tot, ns, dt = run(idx, nmoves)
    out = {"value": round(tot / dt, 3), "unit": "moves/s", "cores": threads, "host_nproc": nproc, "kind": "port",
           "sample": f"{threads} chains (beta, U of chains {idx[0]}..{idx[-1]} of the batch) x {nmoves} moves, "
                     f"cubic2d L={L}, addremove/nn-exchange 0.5/0.5, FP64 Householder+implicit QR per proposal "
                     f"({dt:.1f} s)",
           "eigensolves_per_s": round(ns / dt, 3),
           "single_core_moves_per_s": round(tot1 / dt1, 3), "single_core_eigensolves_per_s": round(ns1 / dt1, 3)}
    # the single-chain workload: cubic2d L=8, U=1, beta=1, mu=0.5, addremove, one core
    rng = np.random.default_rng(8)
    f8 = (rng.random((1, 64)) < 0.5).astype(np.int32)
    t = time.perf_counter()
    tc1, _ = orc.bench_chain_set(orc.KIND_HYPERCUBIC, 2, 8, [1.0], [1.0], [0.5], f8, [7],
                                 [(orc.MOVE_ADDREMOVE, 1.0)], 2000, t=1.0, tp=0.0)
    out["single_chain_cubic2d_L8_moves_per_s"] = round(tc1 / (time.perf_counter() - t), 1)
    # cubic3d L=8 (N=512): U in {2,5} x beta in {2,10}, filling 0.5, one chain per thread
    b3 = np.array([2.0, 10.0] * (threads // 2 + 1))[:threads]
    u3 = np.array([2.0, 2.0, 5.0, 5.0] * (threads // 4 + 1))[:threads]
    f3 = (rng.random((threads, 512)) < 0.5).astype(np.int32)
    t = time.perf_counter()
    t3, n3 = orc.bench_chain_set(orc.KIND_HYPERCUBIC, 3, 8, b3, u3, u3 / 2, f3, np.arange(threads) + 11, moves, 6,
                                 t=1.0, tp=0.0)
    dt3 = time.perf_counter() - t
    out["cubic3d_L8"] = {"moves_per_s": round(t3 / dt3, 2), "eigensolves_per_s": round(n3 / dt3, 2),
                         "cores": threads, "sample": f"{threads} chains x 6 moves ({dt3:.1f} s)"}
    # dOmega = -(logZ_new - logZ_old)/beta of proposals: device trace vs oracle chain
    import fk_mc_amd as fk
    nchk, msteps = 4, 40
    sub = fk.Chains(ctx, f[:nchk], beta[:nchk], U[:nchk], U[:nchk] / 2, U[:nchk] / 2, seeds[:nchk])
    sub.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    tr = sub.sweep(msteps, trace=True)
    sub.close()
    errs, abs_errs = [], []
    for c in range(nchk):
        ref = orc.run_chain(orc.KIND_HYPERCUBIC, 2, L, beta[c], U[c], U[c] / 2, U[c] / 2, int(seeds[c]), f[c],
                            [(orc.MOVE_ADDREMOVE, 0.5), (orc.MOVE_HOP, 0.5)], msteps, t=1.0, tp=0.0,
                            forced=tr["acc"][:, c].astype(np.int32))
        solved = ref["logz"] != 0.0
        lz_old = np.concatenate([[0], ref["logz_cur"][:-1]])
        d_gpu = -(tr["logz"][solved, c] - lz_old[solved]) / beta[c]
        d_ref = -(ref["logz"][solved] - lz_old[solved]) / beta[c]
        om = -ref["logz"][solved] / beta[c]
        if solved[0]:
            d_gpu, d_ref, om = d_gpu[1:], d_ref[1:], om[1:]
        scale = np.maximum(np.abs(d_ref), EPS * np.abs(om))
        errs.append(np.max(np.abs(d_gpu - d_ref) / scale) if len(d_ref) else 0.0)
        abs_errs.append(np.max(np.abs(d_gpu - d_ref)) if len(d_ref) else 0.0)
    out.update({"dOmega_max_rel_err_vs_fp64_oracle": float(max(errs)),
                "dOmega_max_abs_err_vs_fp64_oracle": float(max(abs_errs)),
                "dOmega_err_definition": DOMEGA_ERR_DEF})
    return out


if __name__ == "__main__":
    main()
