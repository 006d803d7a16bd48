"""Train and evaluate the warm-start network on steady states the solver
itself generates.

Training problems are batched-workload problems on a 64^3 grid (one
subdomain, L = 64, h = 16, 8-cell absorber; even indices a spherical well,
odd a Gaussian random field; seeded incidence), energies k ~ U[k_lo, k_hi],
seeds seed_base + i.  They run as 48-problem FP32 batches for a fixed number
of periods of the slowest problem; the harmonic amplitude A = psi^n e^{i tau}
of the final state is the steady state.  The network learns A - A_init from
the potential and the start field (warmstart.features).

Evaluation on held-out seeds: each problem is run to the steady-state
criterion from the plain start and from the warm start (with the
first-period fallback of warmstart.solve_with_warm_start), and the periods
needed are compared.

    python pstd-scalarwavequantumscattering_amd/train_warmstart.py --train 384 --eval 16 \\
        --epochs 20 --out warmstart_weights.npy --summary summary.json
"""
import argparse
import importlib.util
import json
import sys
import time
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent


def load_pkg():
    if "pstd_amd" in sys.modules:
        return sys.modules["pstd_amd"]
    spec = importlib.util.spec_from_file_location("pstd_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pstd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, flush=True)


def make_problem(W, pkg, i, seed_base, k_lo, k_hi):
    k = k_lo + (k_hi - k_lo) * W.SplitMix64(seed_base + i + 1_000_000).uniform()
    return W.batch_problem(i, 64, k, precision=pkg.PSTD_FP32, seed_base=seed_base)


def generate(pkg, W, WS, n, seed_base, k_lo, k_hi, periods, chunk=48):
    """Features (float16) and targets (float32) of n solved problems."""
    X, T = [], []
    for c0 in range(0, n, chunk):
        ids = list(range(c0, min(n, c0 + chunk)))
        probs = [make_problem(W, pkg, i, seed_base, k_lo, k_hi) for i in ids]
        t0 = time.perf_counter()
        b = pkg.PSTDBatch([p for p, _ in probs], [v for _, v in probs])
        a_init = [WS.initial_amplitude(m) for m in b.members]
        nmax = max(m.IterationsToAccumulateSurfaceTerms() for m in b.members)
        total = periods * nmax
        done = 0
        while done < total:
            step = min(4000, total - done)
            b.Update(step)
            b.synchronize()
            done += step
        for (p, V), m, a0 in zip(probs, b.members, a_init):
            X.append(WS.features(p, V, a0).astype(np.float16))
            T.append(WS.target(WS.harmonic_amplitude(m), a0))
        b.close()
        log(f"generated problems {ids[0]}..{ids[-1]}: {total} steps in {time.perf_counter() - t0:.1f}s")
    return X, T


def train(WS, X, T, epochs, lr0, lr1, seed, Xv=None, Tv=None):
    net = WS.WarmStartNet(64, 1, seed=seed)
    rng = np.random.default_rng(seed)
    n = len(X)
    total = epochs * n
    it = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = rng.permutation(n)
        ls = []
        for j in perm:
            lr = lr1 + 0.5 * (lr0 - lr1) * (1 + np.cos(np.pi * it / total))   # cosine schedule
            ls.append(net.train_step(X[j].astype(np.float32)[None], T[j][None], lr))
            it += 1
        msg = f"epoch {ep + 1}/{epochs}: train loss {np.mean(ls):.4e} ({time.perf_counter() - t0:.0f}s)"
        if Xv is not None:
            msg += f", held-out relative error {val_error(net, Xv, Tv):.3f}"
        log(msg)
    return net


def val_error(net, Xv, Tv):
    e = []
    for x, t in zip(Xv, Tv):
        y = net.forward(x.astype(np.float32)[None])[0]
        e.append(np.linalg.norm(y - t) / np.linalg.norm(t))
    return float(np.mean(e))


def evaluate(pkg, W, WS, net, n, seed_base, k_lo, k_hi, tol, max_periods):
    rows = []
    for i in range(n):
        p, V = make_problem(W, pkg, i, seed_base, k_lo, k_hi)
        s = pkg.PSTDQScat(p, V)
        a_init = WS.initial_amplitude(s)
        t0 = time.perf_counter()
        c0, per0, st0, h0 = s.run_to_steady_state(tol=tol, max_periods=max_periods)
        s.synchronize()
        w0 = time.perf_counter() - t0
        a_ss = WS.harmonic_amplitude(s)
        s.close()
        pred = WS.predict(net, p, V, a_init)
        perr = float(np.linalg.norm(pred - a_ss) / np.linalg.norm(a_ss - a_init))
        t1 = time.perf_counter()
        s2, rep = WS.solve_with_warm_start(lambda: pkg.PSTDQScat(p, V), net, p, V, tol, max_periods)
        s2.synchronize()
        w1 = time.perf_counter() - t1
        s2.close()
        row = {"problem": i, "k": float(np.sqrt(2 * p.E0)), "kind": "well" if i % 2 == 0 else "grf",
               "steps_per_period": int(st0 // max(per0, 1)),
               "plain": {"converged": c0, "periods": per0, "seconds": round(w0, 3)},
               "warm": {"converged": rep["converged"], "periods": rep["periods"], "seconds": round(w1, 3),
                        "start": rep["start"]},
               "prediction_rel_error": perr, "first_period_change_network": rep["residual_network"],
               "first_period_change_plain": rep["residual_plain"]}
        rows.append(row)
        log(json.dumps(row))
    both = [r for r in rows if r["plain"]["converged"] and r["warm"]["converged"]]
    summary = {"tol": tol, "max_periods": max_periods, "problems": n,
               "plain_converged": sum(r["plain"]["converged"] for r in rows),
               "warm_converged": sum(r["warm"]["converged"] for r in rows),
               "network_start_used": sum(r["warm"]["start"] == "network" for r in rows),
               "mean_prediction_rel_error": float(np.mean([r["prediction_rel_error"] for r in rows]))}
    if both:
        summary["periods_plain_over_warm"] = float(sum(r["plain"]["periods"] for r in both) /
                                                   sum(r["warm"]["periods"] for r in both))
        summary["seconds_plain_over_warm"] = float(sum(r["plain"]["seconds"] for r in both) /
                                                   sum(r["warm"]["seconds"] for r in both))
        summary["both_converged"] = len(both)
    return rows, summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=384)
    ap.add_argument("--eval", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--periods", type=int, default=40, help="periods of the slowest problem per training batch")
    ap.add_argument("--k-lo", type=float, default=2.4)
    ap.add_argument("--k-hi", type=float, default=3.5)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=1e-5)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--max-periods", type=int, default=60)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="weights file (.npy)")
    ap.add_argument("--weights", default=None, help="skip training and load these weights")
    ap.add_argument("--summary", default=None)
    args = ap.parse_args()
    pkg = load_pkg()
    from pstd_amd import workloads as W, warmstart as WS
    result = {"args": vars(args)}
    if args.weights:
        net = WS.WarmStartNet(64, 1, seed=args.seed)
        net.load(args.weights)
    else:
        X, T = generate(pkg, W, WS, args.train, 100000, args.k_lo, args.k_hi, args.periods)
        Xv, Tv = generate(pkg, W, WS, min(args.eval, 8), 200000, args.k_lo, args.k_hi, args.periods)
        t0 = time.perf_counter()
        net = train(WS, X, T, args.epochs, args.lr, args.lr_end, args.seed, Xv, Tv)
        result["train_seconds"] = time.perf_counter() - t0
        result["heldout_rel_error"] = val_error(net, Xv, Tv)
        del X, T
        if args.out:
            net.save(args.out)
            log(f"saved {args.out}")
    if args.eval > 0:
        rows, summary = evaluate(pkg, W, WS, net, args.eval, 200000, args.k_lo, args.k_hi, args.tol, args.max_periods)
        result["eval"] = rows
        result["summary"] = summary
        log("summary " + json.dumps(summary))
    if args.summary:
        Path(args.summary).write_text(json.dumps(result, indent=1))


if __name__ == "__main__":
    main()
